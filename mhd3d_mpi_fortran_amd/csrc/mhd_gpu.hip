// C-ABI of the MI355X ideal-MHD step (declarations in include/mhd_gpu.h).
//
// One context = one block (the whole domain) on one GPU, one HIP stream, one
// device arena holding the reference-layout state plus the per-stage
// products (face fluxes, edge EMFs).  The step sequences below follow
// step_ab / step_rk2 / step_rk3 (mhd3d3.f90:3602-3792) kernel for kernel.
#include "../../include/mhd_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "mhd_fused.hpp"
#include "mhd_launch.hpp"
#include "mhd_step_kernels.hpp"

using namespace mhd;

namespace {

size_t align_up(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

void mhd_default_params(mhd_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->nx0 = p->ny0 = p->nz0 = 16;
    p->ng = 4;
    p->gam = 5. / 3.;
    p->cfl = .3;
    p->pdmb = 4.;
    p->eps = 1.e-8;
    p->alpha = 2.;
    p->beta = 4.;
    p->reconop = 1;
    p->stepop = 2;
    p->bc[0] = p->bc[1] = p->bc[2] = MHD_BC_EVEN;
    p->precision = MHD_FP64;
    p->surrogate = 0;
    p->device = 0;
}

const char* mhd_error_string(int code) {
    switch (code) {
    case MHD_OK: return "ok";
    case MHD_ERR_HIP: return "HIP runtime error";
    case MHD_ERR_PARAM: return "invalid parameter or size";
    case MHD_ERR_STATE: return "state check failed (rho<0, NaN or p<=0)";
    case MHD_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown error";
    }
}

static int min_ghosts(int reconop) { return reconop == 1 ? 4 : (reconop == 2 ? 2 : 3); }

// ---------------------------------------------------------------------------
// scratch of the entry point that allocates nothing (mhd_reconstruct_noalloc):
// one file-scope device array per GPU, its address looked up once per device
// with that device current.  Host code uses only the looked-up address: the
// array's own name is a host-side stand-in, not a device address.
// ---------------------------------------------------------------------------
constexpr long kScratchDoubles = 1L << 22;   // 32 MiB per device
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kScratchDevices] = {};

// with `device` current: the device address of this GPU's scratch
static double* scratch_on(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[device];
}

int mhd_create(mhd_ctx** out, const mhd_params* p) {
    if (!out || !p) return MHD_ERR_PARAM;
    *out = nullptr;
    if (p->reconop < 1 || p->reconop > 4 || p->stepop < 1 || p->stepop > 3) return MHD_ERR_PARAM;
    if (p->ng < min_ghosts(p->reconop)) return MHD_ERR_PARAM;
    // the ghost fill copies from owned cells only: each axis needs nx0 >= ng
    if (p->nx0 < p->ng || p->ny0 < p->ng || p->nz0 < p->ng) return MHD_ERR_PARAM;
    // index spaces (ghost-inclusive faces and edges) must stay below 2^29, so
    // that every binary64 array stays below 2^32 bytes (32-bit byte offsets
    // in the stencil kernels); about 800^3 cells, which is what the HBM of
    // one GPU holds for the step's arrays
    if ((double)(p->nx0 + 2 * p->ng + 1) * (p->ny0 + 2 * p->ng + 1) * (p->nz0 + 2 * p->ng + 1) >= 536870912.)
        return MHD_ERR_PARAM;
    for (int a = 0; a < 3; ++a)
        if (p->bc[a] != MHD_BC_EVEN && p->bc[a] != MHD_BC_PERIODIC) return MHD_ERR_PARAM;
    if (p->precision != MHD_FP64 && p->precision != MHD_MIXED) return MHD_ERR_PARAM;
    if (p->surrogate && (p->reconop != 1 || p->precision != MHD_MIXED)) return MHD_ERR_UNSUPPORTED;
    if (hipSetDevice(p->device) != hipSuccess) return MHD_ERR_HIP;
    // set-up of the no-allocation entry point's scratch on this device (a
    // failure here surfaces at that entry point, which retries)
    (void)scratch_on(p->device);

    mhd_ctx* c = new mhd_ctx();
    c->p = *p;
    Dims& d = c->d;
    d.nx0 = p->nx0;
    d.ny0 = p->ny0;
    d.nz0 = p->nz0;
    d.ng = p->ng;
    d.nx = p->nx0 + 2 * p->ng;
    d.ny = p->ny0 + 2 * p->ng;
    d.nz = p->nz0 + 2 * p->ng;
    c->ncell = (long)d.nx * d.ny * d.nz;
    c->nbi = (long)(d.nx + 1) * d.ny * d.nz;
    c->nbj = (long)d.nx * (d.ny + 1) * d.nz;
    c->nbk = (long)d.nx * d.ny * (d.nz + 1);
    c->nf[0] = (long)(d.nx0 + 1) * d.ny0 * d.nz0;
    c->nf[1] = (long)d.nx0 * (d.ny0 + 1) * d.nz0;
    c->nf[2] = (long)d.nx0 * d.ny0 * (d.nz0 + 1);
    c->ne[0] = (long)d.nx0 * (d.ny0 + 1) * (d.nz0 + 1);
    c->ne[1] = (long)(d.nx0 + 1) * d.ny0 * (d.nz0 + 1);
    c->ne[2] = (long)(d.nx0 + 1) * (d.ny0 + 1) * d.nz0;
    c->ninterior = (long)d.nx0 * d.ny0 * d.nz0;
    c->emf_fill = kEmfFillThreads;
    // the two-stream stage pays for its fork and join only on blocks whose
    // kernels fill the chip for long; small blocks keep one stream
    c->overlap = c->ninterior >= kOverlapCells;
    if (const char* ev = getenv("MHD_EMF_FILL_THREADS")) c->emf_fill = atol(ev);
    long nmax = c->ncell;
    nmax = c->nbi > nmax ? c->nbi : nmax;
    nmax = c->nbj > nmax ? c->nbj : nmax;
    nmax = c->nbk > nmax ? c->nbk : nmax;
    c->nmax = nmax;
    c->ndtpart = (nmax + kBlock - 1) / kBlock;
    const size_t tf = p->precision == MHD_MIXED ? sizeof(float) : sizeof(double);

    // arena layout
    std::vector<std::pair<void**, size_t>> plan;
    auto cell = [&](double** q) { plan.push_back({(void**)q, align_up(c->ncell * sizeof(double))}); };
    double** cells[] = {&c->rho, &c->vx, &c->vy, &c->vz, &c->pr, &c->bx, &c->by, &c->bz, &c->vf};
    for (auto q : cells) cell(q);
    plan.push_back({(void**)&c->bi, align_up(c->nbi * sizeof(double))});
    plan.push_back({(void**)&c->bj, align_up(c->nbj * sizeof(double))});
    plan.push_back({(void**)&c->bk, align_up(c->nbk * sizeof(double))});
    plan.push_back({(void**)&c->dx, align_up(d.nx * sizeof(double))});
    plan.push_back({(void**)&c->dy, align_up(d.ny * sizeof(double))});
    plan.push_back({(void**)&c->dz, align_up(d.nz * sizeof(double))});
    for (int a = 0; a < 3; ++a) {
        plan.push_back({&c->flux[a], align_up(8 * c->nf[a] * tf)});
        plan.push_back({&c->emf[a], align_up(c->ne[a] * tf)});
    }
    plan.push_back({(void**)&c->dtpart, align_up(c->ndtpart * sizeof(double))});
    plan.push_back({(void**)&c->divpart, align_up(blocks(c->ninterior).x * sizeof(double))});
    plan.push_back({(void**)&c->scal, align_up(8 * sizeof(double))});
    plan.push_back({(void**)&c->flags, align_up(64)});
    const bool mixed = p->precision == MHD_MIXED;
    if (mixed) {
        for (int a = 0; a < 12; ++a) {
            const long n = a < 9 ? c->ncell : (a == 9 ? c->nbi : (a == 10 ? c->nbj : c->nbk));
            plan.push_back({(void**)&c->sh[a], align_up(n * sizeof(float))});
        }
    }
    if (p->stepop == 1) {
        for (int a = 0; a < 11; ++a) {
            const long n = a < 8 ? c->ncell : (a == 8 ? c->nbi : (a == 9 ? c->nbj : c->nbk));
            plan.push_back({(void**)&c->xp[a], align_up(n * sizeof(double))});
            plan.push_back({(void**)&c->xn[a], align_up(n * sizeof(double))});
            if (!mixed) plan.push_back({(void**)&c->xh[a], align_up(n * sizeof(double))});
        }
        if (!mixed) plan.push_back({(void**)&c->vfh, align_up(c->ncell * sizeof(double))});
    } else {
        double** prev[] = {&c->rhop, &c->mxp, &c->myp, &c->mzp, &c->engp};
        for (auto q : prev) cell(q);
        plan.push_back({(void**)&c->bip, align_up(c->nbi * sizeof(double))});
        plan.push_back({(void**)&c->bjp, align_up(c->nbj * sizeof(double))});
        plan.push_back({(void**)&c->bkp, align_up(c->nbk * sizeof(double))});
    }
    size_t total = 0;
    for (auto& e : plan) total += e.second;
    if (hipMalloc(&c->arena, total) != hipSuccess) {
        delete c;
        return MHD_ERR_HIP;
    }
    c->arena_bytes = total;
    size_t off = 0;
    for (auto& e : plan) {
        *e.first = c->arena + off;
        off += e.second;
    }
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
        hipFree(c->arena);
        delete c;
        return MHD_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_emf, hipEventDisableTiming) != hipSuccess) {
        mhd_destroy(c);
        return MHD_ERR_HIP;
    }
    // zero everything once so ghost shells and unused slots hold defined values
    if (hipMemsetAsync(c->arena, 0, total, c->s) != hipSuccess || hipStreamSynchronize(c->s) != hipSuccess) {
        mhd_destroy(c);
        return MHD_ERR_HIP;
    }
    *out = c;
    return MHD_OK;
}

void mhd_destroy(mhd_ctx* c) {
    if (!c) return;
    hipSetDevice(c->p.device);
    if (c->s) hipStreamSynchronize(c->s);
    if (c->s2) hipStreamSynchronize(c->s2);
    for (auto& e : c->evpool) hipEventDestroy(e);
    for (auto& r : c->evs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    surrogate_free(c);
    if (c->metbuf) hipFree(c->metbuf);
    if (c->gmetbuf) hipFree(c->gmetbuf);
    if (c->arena) hipFree(c->arena);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->ev_emf) hipEventDestroy(c->ev_emf);
    if (c->s2) hipStreamDestroy(c->s2);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
}

static int upload_metric_spacing(mhd_ctx* c);
static void ensure_cellb(mhd_ctx* c);

int mhd_set_geometry(mhd_ctx* c, const double* dx, const double* dy, const double* dz) {
    if (!c || !dx || !dy || !dz) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    c->hdx.assign(dx, dx + c->d.nx);
    c->hdy.assign(dy, dy + c->d.ny);
    c->hdz.assign(dz, dz + c->d.nz);
    if (hipMemcpyAsync(c->dx, dx, c->d.nx * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess ||
        hipMemcpyAsync(c->dy, dy, c->d.ny * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess ||
        hipMemcpyAsync(c->dz, dz, c->d.nz * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    c->have_geometry = true;
    // periodic ghost cell g repeats owned cell g +- n0 (1-based); with equal
    // spacings too, a ghost cell's CFL step is its owned twin's
    c->ghost_twins = true;
    const std::vector<double>* hs[3] = {&c->hdx, &c->hdy, &c->hdz};
    const int n0[3] = {c->d.nx0, c->d.ny0, c->d.nz0};
    for (int a = 0; a < 3; ++a) {
        if (c->p.bc[a] != MHD_BC_PERIODIC) c->ghost_twins = false;
        const std::vector<double>& h = *hs[a];
        for (int g = 1; g <= c->d.ng && c->ghost_twins; ++g) {
            const int hi = c->d.ng + n0[a] + g;  // upper ghost g
            if (!(h[g - 1] == h[g + n0[a] - 1]) || !(h[hi - 1] == h[hi - n0[a] - 1])) c->ghost_twins = false;
        }
    }
    if (c->met.on && upload_metric_spacing(c) != MHD_OK) return MHD_ERR_HIP;
    return hipStreamSynchronize(c->s) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

// ---------------------------------------------------------------------------
// affine metric (no counterpart in the reference): x_phys = A x_comp
// ---------------------------------------------------------------------------
static int upload_metric_spacing(mhd_ctx* c) {
    if (!c->have_geometry || !c->dxm) return MHD_OK;
    std::vector<double> h;
    const std::vector<double>* src[3] = {&c->hdx, &c->hdy, &c->hdz};
    double* dst[3] = {c->dxm, c->dym, c->dzm};
    for (int a = 0; a < 3; ++a) {
        h.resize(src[a]->size());
        for (size_t q = 0; q < h.size(); ++q) h[q] = (*src[a])[q] / c->met.g[a];
        if (hipMemcpy(dst[a], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return MHD_ERR_HIP;
    }
    return MHD_OK;
}

// face-normal velocity arrays and the metric spacing (once per context)
static int alloc_metbuf(mhd_ctx* c) {
    if (c->metbuf) return MHD_OK;
    const size_t cb = c->ncell * sizeof(double), fb = c->ncell * sizeof(float);
    const size_t sb = (c->d.nx + c->d.ny + c->d.nz) * sizeof(double);
    const bool ab = c->p.stepop == 1, mx = c->p.precision == MHD_MIXED;
    const size_t total = sb + 3 * cb + (ab && !mx ? 3 * cb : 0) + (mx ? 3 * fb : 0) + 8 * 256;
    if (hipMalloc(&c->metbuf, total) != hipSuccess) {
        c->metbuf = nullptr;
        return MHD_ERR_HIP;
    }
    char* q = c->metbuf;
    auto take = [&](size_t n) {
        char* r = q;
        q += (n + 255) & ~size_t(255);
        return r;
    };
    c->dxm = (double*)take(c->d.nx * sizeof(double));
    c->dym = (double*)take(c->d.ny * sizeof(double));
    c->dzm = (double*)take(c->d.nz * sizeof(double));
    for (int d = 0; d < 3; ++d) c->w64[d] = (double*)take(cb);
    if (ab && !mx)
        for (int d = 0; d < 3; ++d) c->wh64[d] = (double*)take(cb);
    if (mx)
        for (int d = 0; d < 3; ++d) c->w32[d] = (float*)take(fb);
    return MHD_OK;
}

int mhd_set_metric(mhd_ctx* c, const double* A) {
    if (!c || !A) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    // columns a_d = A e_d; cofactor columns C e_d = a_(d+1) x a_(d+2)
    double a[3][3], C[3][3];
    for (int d = 0; d < 3; ++d)
        for (int r = 0; r < 3; ++r) a[d][r] = A[3 * r + d];
    for (int d = 0; d < 3; ++d) {
        const double* p = a[(d + 1) % 3];
        const double* q = a[(d + 2) % 3];
        C[d][0] = p[1] * q[2] - p[2] * q[1];
        C[d][1] = p[2] * q[0] - p[0] * q[2];
        C[d][2] = p[0] * q[1] - p[1] * q[0];
    }
    const double J = a[0][0] * C[0][0] + a[0][1] * C[0][1] + a[0][2] * C[0][2];
    if (!(J > 0.) || !std::isfinite(J)) return MHD_ERR_PARAM;
    Metric m;
    m.on = true;
    m.mode = 1;
    for (int q = 0; q < 9; ++q) m.A[q] = A[q];
    double len[3];
    for (int d = 0; d < 3; ++d) {
        len[d] = std::sqrt(C[d][0] * C[d][0] + C[d][1] * C[d][1] + C[d][2] * C[d][2]);
        for (int r = 0; r < 3; ++r) m.n[d][r] = C[d][r] / len[d];
        m.g[d] = len[d] / J;
        m.ct[d] = 1. / len[d];
    }
    for (int d = 0; d < 3; ++d) m.ef[d] = len[(d + 1) % 3] * len[(d + 2) % 3] / J;
    for (int r = 0; r < 3; ++r)
        for (int d = 0; d < 3; ++d) m.M[3 * r + d] = A[3 * r + d] * m.g[d];
    if (alloc_metbuf(c) != MHD_OK) return MHD_ERR_HIP;
    c->met = m;
    c->dt_fresh = false;
    if (upload_metric_spacing(c) != MHD_OK) return MHD_ERR_HIP;
    return MHD_OK;
}

// general metric from the node coordinates X, Y, Z ((nx+1)(ny+1)(nz+1) each,
// column-major, node (i,j,k) = lower corner of cell (i,j,k), ghost cells
// included).  Face area vectors S = (d1 x d2)/2 of the face diagonals
// (oriented along increasing index), volumes by the divergence theorem with
// face centroids, cell B coefficients of the area-weighted least-squares fit
// of B to the six face normal components,
// and for each owned edge the factor |S1||S2|/V of its local basis (edge
// vector dl and the averaged transverse node differences a1, a2), which
// turns the face-normal velocities and fields into E . dl exactly for a
// uniform state on an affine grid.
long mhd_metric_nodes_host(int nx0, int ny0, int nz0, int ng, const double* X, const double* Y, const double* Z,
                           double* out) {
    if (nx0 < 1 || ny0 < 1 || nz0 < 1 || ng < 1) return -MHD_ERR_PARAM;
    const int nx = nx0 + 2 * ng, ny = ny0 + 2 * ng, nz = nz0 + 2 * ng;
    const long ncell = (long)nx * ny * nz;
    const long nb[3] = {(long)(nx + 1) * ny * nz, (long)nx * (ny + 1) * nz, (long)nx * ny * (nz + 1)};
    const long ne[3] = {(long)nx0 * (ny0 + 1) * (nz0 + 1), (long)(nx0 + 1) * ny0 * (nz0 + 1),
                        (long)(nx0 + 1) * (ny0 + 1) * nz0};
    long total = 31 * ncell;  // 1/V, 18 cell-B coefficients, 9 normals, 3 CFL factors
    for (int f = 0; f < 3; ++f) total += 5 * nb[f] + ne[f];
    if (!out) return total;
    if (!X || !Y || !Z) return -MHD_ERR_PARAM;
    const long nn1 = nx + 1, nn2 = ny + 1;
    auto node = [&](int i, int j, int k, double* p) {  // 1-based node
        const long q = (long)(i - 1) + nn1 * ((long)(j - 1) + nn2 * (long)(k - 1));
        p[0] = X[q];
        p[1] = Y[q];
        p[2] = Z[q];
    };
    auto sub = [](const double* a, const double* b, double* r) {
        for (int q = 0; q < 3; ++q) r[q] = a[q] - b[q];
    };
    auto cross = [](const double* a, const double* b, double* r) {
        r[0] = a[1] * b[2] - a[2] * b[1];
        r[1] = a[2] * b[0] - a[0] * b[2];
        r[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto norm = [&](const double* a) { return std::sqrt(dot(a, a)); };
    // face geometry: area vector and centroid of face (i,j,k) of family f (1-based, ghost-inclusive)
    auto face = [&](int f, int i, int j, int k, double* S, double* xc) {
        double p00[3], p10[3], p01[3], p11[3];  // corners in the face's (a, b) directions
        if (f == 0) {       // a = y, b = z
            node(i, j, k, p00); node(i, j + 1, k, p10); node(i, j, k + 1, p01); node(i, j + 1, k + 1, p11);
        } else if (f == 1) {  // a = z, b = x
            node(i, j, k, p00); node(i, j, k + 1, p10); node(i + 1, j, k, p01); node(i + 1, j, k + 1, p11);
        } else {            // a = x, b = y
            node(i, j, k, p00); node(i + 1, j, k, p10); node(i, j + 1, k, p01); node(i + 1, j + 1, k, p11);
        }
        double d1[3], d2[3];
        sub(p11, p00, d1);
        sub(p01, p10, d2);
        cross(d1, d2, S);
        for (int q = 0; q < 3; ++q) {
            S[q] *= .5;
            xc[q] = .25 * (p00[q] + p10[q] + p01[q] + p11[q]);
        }
    };
    // host arrays
    std::vector<double> nf[3], af[3], iaf[3], ef[3];
    for (int f = 0; f < 3; ++f) {
        nf[f].assign(3 * nb[f], 0.);
        af[f].assign(nb[f], 0.);
        iaf[f].assign(nb[f], 0.);
        ef[f].assign(ne[f], 0.);
    }
    std::vector<double> ivol(ncell), cm(18 * ncell), nc(9 * ncell), ih(3 * ncell);
    const int fe[3][3] = {{nx + 1, ny, nz}, {nx, ny + 1, nz}, {nx, ny, nz + 1}};
    for (int f = 0; f < 3; ++f)
        for (int k = 1; k <= fe[f][2]; ++k)
            for (int j = 1; j <= fe[f][1]; ++j)
                for (int i = 1; i <= fe[f][0]; ++i) {
                    double S[3], xc[3];
                    face(f, i, j, k, S, xc);
                    const long q = (long)(i - 1) + (long)fe[f][0] * ((long)(j - 1) + (long)fe[f][1] * (k - 1));
                    const double a = norm(S);
                    if (!(a > 0.)) return -MHD_ERR_PARAM;
                    af[f][q] = a;
                    iaf[f][q] = 1. / a;
                    for (int r = 0; r < 3; ++r) nf[f][r * nb[f] + q] = S[r] / a;
                }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const long t = (long)(i - 1) + (long)nx * ((long)(j - 1) + (long)ny * (k - 1));
                double S[3][2][3], xf[3][2][3];
                for (int f = 0; f < 3; ++f) {
                    face(f, i, j, k, S[f][0], xf[f][0]);
                    face(f, i + (f == 0), j + (f == 1), k + (f == 2), S[f][1], xf[f][1]);
                }
                double V = 0.;
                for (int f = 0; f < 3; ++f) V += dot(S[f][1], xf[f][1]) - dot(S[f][0], xf[f][0]);
                V /= 3.;
                if (!(V > 0.)) return -MHD_ERR_PARAM;
                ivol[t] = 1. / V;
                // cell B from the six face normal components: the area-weighted
                // least-squares fit B = N^-1 sum_f |S_f| n_f beta_f, N = sum_f
                // |S_f| n_f n_f^T -- exact for a uniform field on any grid (a
                // curved cell's faces are not planar, so the divergence-theorem
                // form (1/V) sum_f (B.S_f)(x_f - x_c) is exact on affine cells
                // only), and on affine cells the same map as that form
                double N[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}}, un[6][3], wf[6];
                for (int f = 0; f < 6; ++f) {
                    const double* Sv = S[f / 2][f % 2];
                    wf[f] = norm(Sv);
                    for (int r = 0; r < 3; ++r) un[f][r] = Sv[r] / wf[f];
                    for (int r = 0; r < 3; ++r)
                        for (int q = 0; q < 3; ++q) N[r][q] += wf[f] * un[f][r] * un[f][q];
                }
                double Ni[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q)
                        Ni[q][r] = N[(r + 1) % 3][(q + 1) % 3] * N[(r + 2) % 3][(q + 2) % 3] -
                                   N[(r + 1) % 3][(q + 2) % 3] * N[(r + 2) % 3][(q + 1) % 3];
                const double detN = N[0][0] * Ni[0][0] + N[0][1] * Ni[1][0] + N[0][2] * Ni[2][0];
                if (!(detN > 0.)) return -MHD_ERR_PARAM;
                for (int f = 0; f < 6; ++f)
                    for (int r = 0; r < 3; ++r)
                        cm[(f * 3 + r) * ncell + t] =
                            wf[f] * (Ni[r][0] * un[f][0] + Ni[r][1] * un[f][1] + Ni[r][2] * un[f][2]) / detN;
                for (int f = 0; f < 3; ++f) {
                    const double a0 = norm(S[f][0]), a1 = norm(S[f][1]);
                    double sn[3];
                    for (int r = 0; r < 3; ++r) sn[r] = S[f][0][r] + S[f][1][r];
                    const double ln = norm(sn);
                    for (int r = 0; r < 3; ++r) nc[(f * 3 + r) * ncell + t] = sn[r] / ln;
                    ih[f * ncell + t] = .5 * (a0 + a1) / V;
                }
            }
    // edges: along e, owned layouts ei (nx0,ny0+1,nz0+1), ej (nx0+1,ny0,nz0+1), ek (nx0+1,ny0+1,nz0)
    for (int e = 0; e < 3; ++e) {
        const int n1 = nx0 + (e != 0), n2 = ny0 + (e != 1), n3 = nz0 + (e != 2);
        const int a = (e + 1) % 3, b = (e + 2) % 3;  // a1 along axis a, a2 along axis b
        for (int k = 1; k <= n3; ++k)
            for (int j = 1; j <= n2; ++j)
                for (int i = 1; i <= n1; ++i) {
                    int I[3] = {ng + i, ng + j, ng + k};
                    double p0[3], p1[3], dl[3], a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
                    node(I[0], I[1], I[2], p0);
                    int J[3] = {I[0], I[1], I[2]};
                    J[e] += 1;
                    node(J[0], J[1], J[2], p1);
                    sub(p1, p0, dl);
                    for (int s2 = 0; s2 < 2; ++s2) {  // the edge's two end nodes
                        int B[3] = {I[0], I[1], I[2]};
                        B[e] += s2;
                        int P[3] = {B[0], B[1], B[2]}, M[3] = {B[0], B[1], B[2]};
                        P[a] += 1;
                        M[a] -= 1;
                        double pp[3], pm[3];
                        node(P[0], P[1], P[2], pp);
                        node(M[0], M[1], M[2], pm);
                        for (int r = 0; r < 3; ++r) a1[r] += .25 * (pp[r] - pm[r]);
                        int Q[3] = {B[0], B[1], B[2]}, N[3] = {B[0], B[1], B[2]};
                        Q[b] += 1;
                        N[b] -= 1;
                        node(Q[0], Q[1], Q[2], pp);
                        node(N[0], N[1], N[2], pm);
                        for (int r = 0; r < 3; ++r) a2[r] += .25 * (pp[r] - pm[r]);
                    }
                    double S1[3], S2[3], t3[3];
                    cross(a2, dl, S1);
                    cross(dl, a1, S2);
                    cross(a2, dl, t3);
                    const double Ve = dot(a1, t3);
                    if (!(Ve > 0.)) return -MHD_ERR_PARAM;
                    ef[e][(long)(i - 1) + (long)n1 * ((long)(j - 1) + (long)n2 * (k - 1))] = norm(S1) * norm(S2) / Ve;
                }
    }
    double* o = out;
    auto emit = [&](const std::vector<double>& v) {
        std::copy(v.begin(), v.end(), o);
        o += v.size();
    };
    for (int f = 0; f < 3; ++f) {
        emit(nf[f]);
        emit(af[f]);
        emit(iaf[f]);
    }
    for (int e = 0; e < 3; ++e) emit(ef[e]);
    emit(ivol);
    emit(cm);
    emit(nc);
    emit(ih);
    return total;
}

int mhd_set_metric_nodes(mhd_ctx* c, const double* X, const double* Y, const double* Z) {
    if (!c || !X || !Y || !Z || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    const Dims& d = c->d;
    const long total = mhd_metric_nodes_host(d.nx0, d.ny0, d.nz0, d.ng, X, Y, Z, nullptr);
    std::vector<double> host((size_t)total);
    const long r = mhd_metric_nodes_host(d.nx0, d.ny0, d.nz0, d.ng, X, Y, Z, host.data());
    if (r < 0) return (int)-r;
    const long nb[3] = {c->nbi, c->nbj, c->nbk};
    const long ncell = c->ncell;
    // device buffer: the host arrays in the same order, each 256-byte aligned
    if (alloc_metbuf(c) != MHD_OK) return MHD_ERR_HIP;
    auto al = [](size_t n) { return (n * sizeof(double) + 255) & ~size_t(255); };
    size_t bytes = 0;
    for (int f = 0; f < 3; ++f) bytes += al(3 * nb[f]) + 2 * al(nb[f]) + al(c->ne[f]);
    bytes += al(ncell) + al(18 * ncell) + al(9 * ncell) + al(3 * ncell);
    if (c->gmetbuf) hipFree(c->gmetbuf);
    c->gmetbuf = nullptr;
    if (hipMalloc(&c->gmetbuf, bytes) != hipSuccess) {
        c->gmetbuf = nullptr;
        return MHD_ERR_HIP;
    }
    char* q = c->gmetbuf;
    const double* h = host.data();
    bool ok = true;
    auto put = [&](size_t n) -> const double* {
        double* dst = (double*)q;
        q += al(n);
        ok = ok && hipMemcpy(dst, h, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
        h += n;
        return dst;
    };
    GMet g{};
    for (int f = 0; f < 3; ++f) {
        g.nf[f] = put(3 * nb[f]);
        g.af[f] = put(nb[f]);
        g.iaf[f] = put(nb[f]);
        g.nb[f] = nb[f];
    }
    for (int f = 0; f < 3; ++f) g.ef[f] = put(c->ne[f]);
    g.ivol = put(ncell);
    g.cm = put(18 * ncell);
    g.nc = put(9 * ncell);
    g.ih = put(3 * ncell);
    g.ncell = ncell;
    if (!ok) return MHD_ERR_HIP;
    Metric m;
    m.on = true;
    m.mode = 2;
    c->met = m;
    c->gm = g;
    c->dt_fresh = false;
    if (upload_metric_spacing(c) != MHD_OK) return MHD_ERR_HIP;
    return MHD_OK;
}

static const double* sdx(const mhd_ctx* c) { return c->met.on ? c->dxm : c->dx; }
static const double* sdy(const mhd_ctx* c) { return c->met.on ? c->dym : c->dy; }
static const double* sdz(const mhd_ctx* c) { return c->met.on ? c->dzm : c->dz; }

int mhd_upload_state(mhd_ctx* c, const double* rho, const double* vx, const double* vy, const double* vz,
                     const double* p, const double* bi, const double* bj, const double* bk) {
    if (!c || !rho || !vx || !vy || !vz || !p || !bi || !bj || !bk) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    const size_t cb = c->ncell * sizeof(double);
    const void* src[8] = {rho, vx, vy, vz, p, bi, bj, bk};
    void* dst[8] = {c->rho, c->vx, c->vy, c->vz, c->pr, c->bi, c->bj, c->bk};
    const size_t nb[8] = {cb, cb, cb, cb, cb, c->nbi * sizeof(double), c->nbj * sizeof(double),
                          c->nbk * sizeof(double)};
    for (int a = 0; a < 8; ++a)
        if (hipMemcpyAsync(dst[a], src[a], nb[a], hipMemcpyHostToDevice, c->s) != hipSuccess) return MHD_ERR_HIP;
    // cell B of the new faces is formed where it is read (get_dt, downloads)
    c->dt_fresh = false;
    c->cellb_stale = true;
    c->pred_valid = false;
    return hipStreamSynchronize(c->s) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

int mhd_download_state(mhd_ctx* c, double* rho, double* vx, double* vy, double* vz, double* p, double* bi,
                       double* bj, double* bk, double* bx, double* by, double* bz) {
    if (!c) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    if (bx || by || bz) ensure_cellb(c);
    const size_t cb = c->ncell * sizeof(double);
    void* dst[11] = {rho, vx, vy, vz, p, bi, bj, bk, bx, by, bz};
    const void* src[11] = {c->rho, c->vx, c->vy, c->vz, c->pr, c->bi, c->bj, c->bk, c->bx, c->by, c->bz};
    const size_t nb[11] = {cb, cb, cb, cb, cb, c->nbi * sizeof(double), c->nbj * sizeof(double),
                           c->nbk * sizeof(double), cb, cb, cb};
    for (int a = 0; a < 11; ++a)
        if (dst[a] && hipMemcpyAsync(dst[a], src[a], nb[a], hipMemcpyDeviceToHost, c->s) != hipSuccess)
            return MHD_ERR_HIP;
    return hipStreamSynchronize(c->s) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

// ---------------------------------------------------------------------------
// building blocks
// ---------------------------------------------------------------------------
static Prim<double> cur_state(mhd_ctx* c) {
    return Prim<double>{c->rho, c->vx, c->vy, c->vz, c->pr, c->bx, c->by, c->bz};
}
static Src<double> cur_src(mhd_ctx* c) {
    Src<double> s{c->rho, c->vx, c->vy, c->vz, c->pr, c->bx, c->by, c->bz, c->vf, c->bi, c->bj, c->bk, {}};
    for (int d = 0; d < 3; ++d) s.w[d] = c->w64[d];
    return s;
}
static Src<double> h_src(mhd_ctx* c) {
    Src<double> s{c->xh[0], c->xh[1], c->xh[2], c->xh[3], c->xh[4], c->xh[5],
                  c->xh[6], c->xh[7], c->vfh,   c->xh[8], c->xh[9], c->xh[10], {}};
    for (int d = 0; d < 3; ++d) s.w[d] = c->wh64[d];
    return s;
}
static Src<float> sh_src(mhd_ctx* c) {
    float* const* h = c->sh;
    Src<float> s{h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], h[8], h[9], h[10], h[11], {}};
    for (int d = 0; d < 3; ++d) s.w[d] = c->w32[d];
    return s;
}
// the general metric's device arrays, or an empty GMet (ivol null) for the others
static GMet gmet(const mhd_ctx* c) { return c->met.mode == 2 ? c->gm : GMet{}; }
template <typename TH>
static Proj<TH> proj(const mhd_ctx* c, TH* const* w) {
    Proj<TH> pj{};
    pj.on = c->met.on ? 1 : 0;
    for (int q = 0; q < 9; ++q) pj.n[q] = c->met.n[q / 3][q % 3];
    for (int d = 0; d < 3; ++d) pj.w[d] = w[d];
    pj.nc = c->met.mode == 2 ? c->gm.nc : nullptr;
    pj.ncell = c->ncell;
    return pj;
}
static BMet bmet(mhd_ctx* c, bool w64, bool w32) {
    BMet m{};
    m.on = c->met.on ? 1 : 0;
    for (int q = 0; q < 9; ++q) {
        m.M[q] = c->met.M[q];
        m.n[q] = c->met.n[q / 3][q % 3];
    }
    for (int d = 0; d < 3; ++d) {
        m.w[d] = w64 ? c->w64[d] : nullptr;
        m.ws[d] = w32 ? c->w32[d] : nullptr;
    }
    m.g = gmet(c);
    return m;
}
static bool mixed(const mhd_ctx* c) { return c->p.precision == MHD_MIXED; }
// the mixed Adams-Bashforth update takes the CFL partials of the new state
// itself (instead of a get_bxyz pass over the block) where every ghost cell
// repeats an owned cell
static bool fuse_dt(const mhd_ctx* c) {
    return mixed(c) && c->p.stepop == 1 && c->met.mode != 2 && c->ghost_twins;
}

// the 11 state arrays of the current state (AB rotation order)
static double** state_slot(mhd_ctx* c, int a) {
    double** t[11] = {&c->rho, &c->vx, &c->vy, &c->vz, &c->pr, &c->bx, &c->by, &c->bz, &c->bi, &c->bj, &c->bk};
    return t[a];
}

static void do_fill(mhd_ctx* c) {
    FillSet fs;
    const Dims& d = c->d;
    double* arr[8] = {c->rho, c->pr, c->vx, c->vy, c->vz, c->bi, c->bj, c->bk};
    const int face[8] = {0, 0, 0, 0, 0, 1, 2, 3};
    long most = 0;
    const int g2 = 2 * d.ng;
    for (int a = 0; a < 8; ++a) {
        FillArr& A = fs.arr[a];
        A.a = arr[a];
        A.n[0] = d.nx + (face[a] == 1);
        A.n[1] = d.ny + (face[a] == 2);
        A.n[2] = d.nz + (face[a] == 3);
        A.face = face[a];
        const long o0 = d.nx0 + (face[a] == 1), o1 = d.ny0 + (face[a] == 2);
        const long cnt = (long)g2 * A.n[1] * A.n[2] + o0 * g2 * A.n[2] + o0 * o1 * g2;
        most = cnt > most ? cnt : most;
    }
    fs.narr = 8;
    fs.n0[0] = d.nx0;
    fs.n0[1] = d.ny0;
    fs.n0[2] = d.nz0;
    for (int a = 0; a < 3; ++a) fs.bc[a] = c->p.bc[a];
    dim3 grid((unsigned)((most + kBlock - 1) / kBlock), 8);
    launch(c, MHD_K_FILL, grid, dim3(kBlock), k_fill, fs, d.ng);
    c->dt_fresh = false;
}

// get_bxyz on the current state, plus (optionally) vf, the fp32 stencil
// copies and the CFL partials of the next get_dt
static void do_bxyz(mhd_ctx* c, bool want_vf, bool want_shadow, bool want_dt, bool store_cellb = true) {
    Shadow sh{};
    if (want_shadow && mixed(c))
        sh = Shadow{c->sh[0], c->sh[1], c->sh[2], c->sh[3], c->sh[4],  c->sh[5],
                    c->sh[6], c->sh[7], c->sh[8], c->sh[9], c->sh[10], c->sh[11]};
    // face-normal velocities where the EMF will read them: binary64 with vf,
    // fp32 with the stencil copies
    const BMet m = bmet(c, want_vf && !mixed(c), want_shadow && mixed(c));
    auto kb = c->met.mode == 2 ? k_bxyz<true> : k_bxyz<false>;
    launch(c, MHD_K_BXYZ, blocks(c->nmax), dim3(kBlock), kb, c->d, c->p.gam, c->p.cfl,
           CFaces{c->bi, c->bj, c->bk}, CCells{c->rho, c->vx, c->vy, c->vz, c->pr},
           store_cellb ? c->bx : (double*)nullptr, c->by, c->bz, want_vf ? c->vf : (double*)nullptr, sh, m,
           sdx(c), sdy(c), sdz(c), want_dt ? c->dtpart : (double*)nullptr, c->ncell, c->nbi, c->nbj, c->nbk);
    c->dt_fresh = want_dt;
    c->ndt_valid = blocks(c->nmax).x;
    c->cellb_stale = !store_cellb;
    // mixed precision keeps the extrapolated state in the fp32 copies
    if (want_shadow && mixed(c)) c->pred_valid = false;
}

// get_dt on the current state; the minimum lands in scal[SC_DTCFL] and, when
// set_tau, also in scal[SC_TAU] (Adams-Bashforth takes tau from the CFL)
static void do_dt(mhd_ctx* c, bool set_tau) {
    // where cell B of the current state is not stored (mixed Adams-Bashforth
    // steps, uploads), get_bxyz forms it and the CFL partials in one pass
    if (!c->dt_fresh && c->cellb_stale) do_bxyz(c, false, false, true);
    if (!c->dt_fresh) {
        launch(c, MHD_K_DT, blocks(c->ncell), dim3(kBlock), k_dt_partial, c->d, c->p.gam, c->p.cfl, cur_state(c),
               sdx(c), sdy(c), sdz(c), gmet(c), c->dtpart, c->ncell);
        launch(c, MHD_K_DT, dim3(1), dim3(1024), k_dt_final, (const double*)c->dtpart, blocks(c->ncell).x * 1L,
               c->scal + SC_DTCFL, set_tau ? c->scal + SC_TAU : (double*)nullptr);
        return;
    }
    launch(c, MHD_K_DT, dim3(1), dim3(1024), k_dt_final, (const double*)c->dtpart, c->ndt_valid, c->scal + SC_DTCFL,
           set_tau ? c->scal + SC_TAU : (double*)nullptr);
}

static void do_prev_rk(mhd_ctx* c) {
    launch(c, MHD_K_PREV, blocks(c->nmax), dim3(kBlock), k_prev_rk, c->p.gam - 1., cur_state(c), c->rhop, c->mxp,
           c->myp, c->mzp, c->engp, (const double*)c->bi, (const double*)c->bj, (const double*)c->bk, c->bip,
           c->bjp, c->bkp, c->ncell, c->nbi, c->nbj, c->nbk);
}

// conserved update from the current state into `out` (the current state
// itself for RK, the next-state buffers for AB)
template <typename TF, int NST, int ST>
static void do_cells_t(mhd_ctx* c, double* const* out, bool fused_ok = true) {
    CCells s{c->rho, c->vx, c->vy, c->vz, c->pr};
    Cells so{out[0], out[1], out[2], out[3], out[4]};
    PrevRK pv{c->rhop, c->mxp, c->myp, c->mzp, c->engp};
    FluxSet<TF> F{(const TF*)c->flux[0], (const TF*)c->flux[1], (const TF*)c->flux[2], c->nf[0], c->nf[1], c->nf[2]};
    const bool gen = c->met.mode == 2;
    if constexpr (NST == 0 && std::is_same_v<TF, float>) {
        if (fused_ok && fuse_dt(c)) {  // the CFL partials of the new state too (CT has run)
            UpdDt u{};
            u.bi = out[8];
            u.bj = out[9];
            u.bk = out[10];
            u.on = c->met.on ? 1 : 0;
            for (int q = 0; q < 9; ++q) u.M[q] = c->met.M[q];
            u.gam = c->p.gam;
            u.cfl = c->p.cfl;
            u.part = c->dtpart;
            launch(c, MHD_K_UPDATE, blocks(c->ninterior), dim3(kBlock), k_update<0, 0, float, false, true>, c->d,
                   c->p.gam - 1., (const double*)(c->scal + SC_TAU), s, so, pv, F, sdx(c), sdy(c), sdz(c), gmet(c),
                   c->ninterior, u);
            return;
        }
    }
    auto ku = gen ? k_update<NST, ST, TF, true> : k_update<NST, ST, TF, false>;
    launch(c, MHD_K_UPDATE, blocks(c->ninterior), dim3(kBlock), ku, c->d, c->p.gam - 1.,
           (const double*)(c->scal + SC_TAU), s, so, pv, F, sdx(c), sdy(c), sdz(c), gmet(c), c->ninterior, UpdDt{});
}

// CT face update from the stage's EMFs into `out`
template <typename TF, int NST, int ST>
static void do_ct_t(mhd_ctx* c, double* const* out) {
    CFaces b{c->bi, c->bj, c->bk};
    Faces bo{out[8], out[9], out[10]};
    CFaces bp{c->bip, c->bjp, c->bkp};
    const long nct = (long)(c->d.nx0 + 1) * c->d.ny0 * c->d.nz0 + (long)c->d.nx0 * (c->d.ny0 + 1) * c->d.nz0 +
                     (long)c->d.nx0 * c->d.ny0 * (c->d.nz0 + 1);
    const CtFactor ctf{{c->met.ct[0], c->met.ct[1], c->met.ct[2]}};
    auto kc = c->met.mode == 2 ? k_ct<NST, ST, TF, true> : k_ct<NST, ST, TF, false>;
    launch(c, MHD_K_CT, blocks(nct), dim3(kBlock), kc, c->d, (const double*)(c->scal + SC_TAU), ctf, b,
           bo, bp, (const TF*)c->emf[0], (const TF*)c->emf[1], (const TF*)c->emf[2], (const double*)c->dx,
           (const double*)c->dy, (const double*)c->dz, gmet(c));
}

// One stage's stencil work and update on two streams.  The edge EMF and,
// behind it, the CT face update run on the second stream; the face flux and
// the cell update on the main stream.  The EMF reads the stage's
// velocities, faces and vf and writes the EMF arrays; CT reads those and the
// faces and writes the faces of `out`; the flux reads the stage's cell
// primitives and writes the flux arrays; the update reads the fluxes and
// cells and writes the cells of `out`.  So the only ordering needed across
// the streams: the update (which may overwrite the cell velocities the
// binary64 EMF reads) waits for the EMF, and the main stream waits for CT
// before the stage ends.  The EMF kernels wait on latency (few waves per
// SIMD) and CT on HBM, the flux on VALU issue: together they fill the CUs
// better than one after the other.
template <int NST, int ST, typename E, typename F>
static void stage_overlapped(mhd_ctx* c, E emf, F flux, double* const* out) {
    auto cells = [&] {
        if (mixed(c)) do_cells_t<float, NST, ST>(c, out);
        else do_cells_t<double, NST, ST>(c, out);
    };
    auto ct = [&] {
        if (mixed(c)) do_ct_t<float, NST, ST>(c, out);
        else do_ct_t<double, NST, ST>(c, out);
    };
    // an update that takes the CFL partials of the new state reads CT's faces
    const bool after_ct = NST == 0 && fuse_dt(c);
    const bool ok = c->overlap && hipEventRecord(c->ev_fork, c->s) == hipSuccess &&
                    hipStreamWaitEvent(c->s2, c->ev_fork, 0) == hipSuccess;
    if (!ok) {   // serial
        emf();
        flux();
        if (after_ct) ct();
        cells();
        if (!after_ct) ct();
        return;
    }
    c->cur = c->s2;
    emf();
    hipEventRecord(c->ev_emf, c->s2);
    ct();
    hipEventRecord(c->ev_join, c->s2);
    c->cur = nullptr;
    flux();
    hipStreamWaitEvent(c->s, after_ct ? c->ev_join : c->ev_emf, 0);
    cells();
    if (!after_ct) hipStreamWaitEvent(c->s, c->ev_join, 0);
}

// the fused face flux + update (mhd_fused.hip) takes the mixed-precision
// Adams-Bashforth stage with plain PDM on the tensor and affine grids
static bool can_fuse(const mhd_ctx* c) {
    // (its CFL partials, one per workgroup, fit the partials buffer: a
    // workgroup covers at most 64 x 4 x 32 cells, the buffer holds one slot
    // per 256 ghost-inclusive cells)
    return c->fused && mixed(c) && c->p.stepop == 1 && c->p.reconop == 1 && !c->p.surrogate && c->met.mode != 2 &&
           (long)fused_blocks(c->d) <= c->ndtpart;
}

// Adams-Bashforth stage with the fused flux + update: the edge EMF and CT on
// the second stream, the fused kernel on the main one (after CT when it
// takes the CFL partials of CT's new faces)
static void stage_fused(mhd_ctx* c, double* const* out) {
    const bool dt = fuse_dt(c);
    FusedArgs a{};
    a.d = c->d;
    a.c = make_coef<float>(c);
    const Src<float> sh = sh_src(c);
    a.q = Prim<float>{sh.rho, sh.vx, sh.vy, sh.vz, sh.p, sh.bx, sh.by, sh.bz};
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) a.n[r][q] = float(c->met.n[r][q]);
    a.rho = c->rho;
    a.vx = c->vx;
    a.vy = c->vy;
    a.vz = c->vz;
    a.p = c->pr;
    a.orho = out[0];
    a.ovx = out[1];
    a.ovy = out[2];
    a.ovz = out[3];
    a.op = out[4];
    a.dx = sdx(c);
    a.dy = sdy(c);
    a.dz = sdz(c);
    a.tau_p = c->scal + SC_TAU;
    a.gm1 = c->p.gam - 1.;
    a.bi = out[8];
    a.bj = out[9];
    a.bk = out[10];
    for (int q = 0; q < 9; ++q) a.M[q] = c->met.M[q];
    a.met_on = c->met.on ? 1 : 0;
    a.gam = c->p.gam;
    a.cfl = c->p.cfl;
    a.part = c->dtpart;
    const bool ok = c->overlap && hipEventRecord(c->ev_fork, c->s) == hipSuccess &&
                    hipStreamWaitEvent(c->s2, c->ev_fork, 0) == hipSuccess;
    if (!ok) {
        launch_emf_fp32(c, sh);
        do_ct_t<float, 0, 0>(c, out);
        launch_flux_update(c, a, dt);
        return;
    }
    c->cur = c->s2;
    launch_emf_fp32(c, sh);
    do_ct_t<float, 0, 0>(c, out);
    hipEventRecord(c->ev_join, c->s2);
    c->cur = nullptr;
    if (dt) hipStreamWaitEvent(c->s, c->ev_join, 0);
    launch_flux_update(c, a, dt);
    if (!dt) hipStreamWaitEvent(c->s, c->ev_join, 0);
}

template <int NST, int ST>
static void do_stage(mhd_ctx* c, Src<double> s64, double* const* out) {
    if (mixed(c)) {
        stage_overlapped<NST, ST>(c, [&] { launch_emf_fp32(c, sh_src(c)); }, [&] {
            if (c->p.surrogate) launch_flux_surrogate(c, sh_src(c));
            else launch_flux_fp32(c, sh_src(c));
        }, out);
    } else {
        stage_overlapped<NST, ST>(c, [&] { launch_emf_fp64(c, s64); }, [&] { launch_flux_fp64(c, s64); }, out);
    }
}

static int sync_status(mhd_ctx* c) {
    if (c->launch_error != hipSuccess) {
        c->launch_error = hipSuccess;
        return MHD_ERR_HIP;
    }
    return hipStreamSynchronize(c->s) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

// ---------------------------------------------------------------------------
// public per-routine entry points
// ---------------------------------------------------------------------------
int mhd_fill_ghosts(mhd_ctx* c) {
    if (!c) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    do_fill(c);
    return sync_status(c);
}

int mhd_get_bxyz(mhd_ctx* c) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    do_bxyz(c, true, true, true);
    return sync_status(c);
}

int mhd_get_dt(mhd_ctx* c, double* tau_cfl) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    do_dt(c, true);   // get_dt sets the module's tau (mhd3d3.f90:1629)
    double h = 0.;
    if (hipMemcpyAsync(&h, c->scal + SC_DTCFL, sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    int rc = sync_status(c);
    if (tau_cfl) *tau_cfl = h;
    return rc;
}

// which_state: the state the reference passes as the routines' array
// arguments -- the current one (step_rk2/rk3, mhd3d3.f90:3665-3667) or the
// Adams-Bashforth extrapolation of mhd_predict_ab (step_ab, 3618-3620)
static int check_which(const mhd_ctx* c, int which) {
    if (which == MHD_STATE_CURRENT) return MHD_OK;
    if (which != MHD_STATE_PREDICTED) return MHD_ERR_PARAM;
    return c->p.stepop == 1 && c->pred_valid ? MHD_OK : MHD_ERR_PARAM;
}

int mhd_flux_stress(mhd_ctx* c, int which_state) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    if (int rc = check_which(c, which_state)) return rc;
    hipSetDevice(c->p.device);
    const bool pred = which_state == MHD_STATE_PREDICTED;
    if (mixed(c)) {
        // the extrapolated state already sits in the fp32 copies
        if (!pred) do_bxyz(c, false, true, false);
        if (c->p.surrogate) launch_flux_surrogate(c, sh_src(c));
        else launch_flux_fp32(c, sh_src(c));
    } else {
        if (!pred) ensure_cellb(c);
        launch_flux_fp64(c, pred ? h_src(c) : cur_src(c));
    }
    return sync_status(c);
}

int mhd_eijk_lfm(mhd_ctx* c, int which_state) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    if (int rc = check_which(c, which_state)) return rc;
    hipSetDevice(c->p.device);
    if (mixed(c)) {
        if (which_state != MHD_STATE_PREDICTED) do_bxyz(c, false, true, false);
        launch_emf_fp32(c, sh_src(c));
    } else if (which_state == MHD_STATE_PREDICTED) {
        launch_emf_fp64(c, h_src(c));   // vf and face-normal velocities written by the predictor
    } else {
        ensure_cellb(c);
        if (c->met.on) do_bxyz(c, true, false, false);   // vf and the face-normal velocities
        else launch(c, MHD_K_BXYZ, blocks(c->ncell), dim3(kBlock), k_vf, cur_state(c), c->vf, c->ncell);
        launch_emf_fp64(c, cur_src(c));
    }
    return sync_status(c);
}

// get_eijk_Yee (mhd3d3.f90:1654-1752).  The reference builds the twelve
// split fields over the whole block, then for each edge direction reconstructs
// four of them along one transverse axis (left states of pp and pn, right
// states of nn and np, untruncated directions 4-6), those four along the other
// transverse axis (left of pp and np, right of nn and pn) and sums the
// quadrants at the owned edges:
//   ei: y (dir 5) on (nx, ny0+1, nz), then z (dir 6) on (nx, ny0+1, nz0+1)
//   ej: z (dir 6) on (nx, ny, nz0+1), then x (dir 4) on (nx0+1, ny, nz0+1)
//   ek: x (dir 4) on (nx0+1, ny, nz), then y (dir 5) on (nx0+1, ny0+1, nz)
// Here the split is one kernel, each reconstruction is k_recon (the
// reconstruct_3d pinned bit for bit) and the sums a last kernel; the working
// arrays are allocated for the call.
static int eijk_yee(mhd_ctx* c, Prim<double> q) {
    const Dims& d = c->d;
    const int nx = d.nx, ny = d.ny, nz = d.nz, nx0 = d.nx0, ny0 = d.ny0, nz0 = d.nz0;
    const long ncell = (long)nx * ny * nz;
    struct Pass {
        int dir1, l1[3], dir2, l2[3], ax;
    };
    const Pass ps[3] = {{5, {nx, ny0 + 1, nz}, 6, {nx, ny0 + 1, nz0 + 1}, 0},
                        {6, {nx, ny, nz0 + 1}, 4, {nx0 + 1, ny, nz0 + 1}, 1},
                        {4, {nx0 + 1, ny, nz}, 5, {nx0 + 1, ny0 + 1, nz}, 2}};
    long nq = 0, ne = 0;
    for (const Pass& p : ps) {
        nq = std::max(nq, (long)p.l1[0] * p.l1[1] * p.l1[2]);
        ne = std::max(ne, (long)p.l2[0] * p.l2[1] * p.l2[2]);
    }
    // 12 split fields, 4 face states, 4 edge states, one discarded side
    const long total = 12 * ncell + 4 * nq + 4 * ne + std::max(nq, ne);
    double* buf = nullptr;
    if (hipMalloc(&buf, (size_t)total * sizeof(double)) != hipSuccess) return MHD_ERR_HIP;
    double* split[12];
    for (int a = 0; a < 12; ++a) split[a] = buf + a * ncell;
    double* fq[4];
    double* fe[4];
    for (int a = 0; a < 4; ++a) {
        fq[a] = buf + 12 * ncell + a * nq;
        fe[a] = buf + 12 * ncell + 4 * nq + a * ne;
    }
    double* waste = buf + 12 * ncell + 4 * nq + 4 * ne;
    launch_yee_split(c, ncell, q, split);
    const int fn[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        const Pass& p = ps[a];
        double* const* e = split + 4 * a;
        // face states: left of pp and pn, right of nn and np
        launch_recon(c, p.dir1, e[YEE_PP], fn[0], fn[1], p.l1[0], p.l1[1], p.l1[2], fq[YEE_PP], waste);
        launch_recon(c, p.dir1, e[YEE_PN], fn[0], fn[1], p.l1[0], p.l1[1], p.l1[2], fq[YEE_PN], waste);
        launch_recon(c, p.dir1, e[YEE_NN], fn[0], fn[1], p.l1[0], p.l1[1], p.l1[2], waste, fq[YEE_NN]);
        launch_recon(c, p.dir1, e[YEE_NP], fn[0], fn[1], p.l1[0], p.l1[1], p.l1[2], waste, fq[YEE_NP]);
        // edge states: left of pp and np, right of nn and pn
        launch_recon(c, p.dir2, fq[YEE_PP], p.l1[0], p.l1[1], p.l2[0], p.l2[1], p.l2[2], fe[0], waste);
        launch_recon(c, p.dir2, fq[YEE_NP], p.l1[0], p.l1[1], p.l2[0], p.l2[1], p.l2[2], fe[1], waste);
        launch_recon(c, p.dir2, fq[YEE_PN], p.l1[0], p.l1[1], p.l2[0], p.l2[1], p.l2[2], waste, fe[2]);
        launch_recon(c, p.dir2, fq[YEE_NN], p.l1[0], p.l1[1], p.l2[0], p.l2[1], p.l2[2], waste, fe[3]);
        // epp + enp + epn + enn
        const int o[3] = {a == 0 ? nx0 : nx0 + 1, a == 1 ? ny0 : ny0 + 1, a == 2 ? nz0 : nz0 + 1};
        const double* e4[4] = {fe[0], fe[1], fe[2], fe[3]};
        launch_yee_sum(c, o[0], o[1], o[2], p.ax, p.l2[0], p.l2[1], e4, (double*)c->emf[a]);
    }
    const int rc = sync_status(c);
    hipFree(buf);
    return rc;
}

int mhd_eijk_yee(mhd_ctx* c, int which_state) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    if (mixed(c) || c->met.on) return MHD_ERR_UNSUPPORTED;
    if (int rc = check_which(c, which_state)) return rc;
    hipSetDevice(c->p.device);
    if (which_state == MHD_STATE_PREDICTED) {
        const Src<double> h = h_src(c);
        return eijk_yee(c, Prim<double>{h.rho, h.vx, h.vy, h.vz, h.p, h.bx, h.by, h.bz});
    }
    ensure_cellb(c);
    return eijk_yee(c, cur_state(c));
}

int mhd_flux_stress_old(mhd_ctx* c) { return mhd_flux_stress(c, MHD_STATE_PREDICTED); }

// vf and fp32 copies are kept only where the next step reads them: RK reads
// the current state (vf in binary64, the copies in mixed precision), AB
// reads the extrapolated state written by the predictor
static void end_of_stage(mhd_ctx* c, bool in_step = false) {
    const bool rk = c->p.stepop != 1;
    // mixed Adams-Bashforth steps take the extrapolated cell B from the faces
    // (k_predict_ab_mixed), so the step itself leaves cell B implicit
    do_bxyz(c, rk && !mixed(c), rk && mixed(c), true, !(in_step && !rk && mixed(c)));
}
// cell B of the current state where a caller reads it
static void ensure_cellb(mhd_ctx* c) {
    if (c->cellb_stale) do_bxyz(c, false, false, false);
}

int mhd_initialize(mhd_ctx* c, double* tau) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    const double t = tau ? *tau : 0.;
    c->tau0_host = t;
    c->pred_valid = false;
    launch(c, MHD_K_PREV, dim3(1), dim3(64), k_set2, c->scal + SC_TAU, t, t);
    do_fill(c);
    end_of_stage(c);
    double out = t;
    if (c->p.stepop == 1) {
        do_dt(c, true);
        // hard-coded Adams-Bashforth start (mhd3d3.f90:3358-3373)
        out = 3.e-4;
        launch(c, MHD_K_PREV, dim3(1), dim3(64), k_set2, c->scal + SC_TAU, out, out);
        for (int a = 0; a < 11; ++a) {
            const long n = a < 8 ? c->ncell : (a == 8 ? c->nbi : (a == 9 ? c->nbj : c->nbk));
            if (hipMemcpyAsync(c->xp[a], *state_slot(c, a), n * sizeof(double), hipMemcpyDeviceToDevice, c->s) !=
                hipSuccess)
                return MHD_ERR_HIP;
        }
    }
    if (tau) *tau = out;
    return sync_status(c);
}

// step_rk2 / step_rk3 (mhd3d3.f90:3648-3792), in place
template <int NST>
static void step_rk(mhd_ctx* c) {
    do_dt(c, false);
    launch(c, MHD_K_PREV, dim3(1), dim3(64), k_set2, c->scal + SC_TAU, c->tau0_host, c->tau0_host);
    do_prev_rk(c);
    double* self[11];
    for (int a = 0; a < 11; ++a) self[a] = *state_slot(c, a);
    for (int st = 1; st <= NST; ++st) {
        if (st == 1) do_stage<NST, 1>(c, cur_src(c), self);
        else if (st == 2) do_stage<NST, 2>(c, cur_src(c), self);
        else do_stage<NST, 3>(c, cur_src(c), self);
        do_fill(c);
        end_of_stage(c);
    }
}

// step_ab (mhd3d3.f90:3602-3644): the new state goes to the next-state
// buffers, then (previous, current, next) rotate -- the reference's
// advance_previous copy x_p = x without moving data
// advance_predictor_ab (mhd3d3.f90:3821-3840) up to its advance_previous:
// the extrapolated state into xh (binary64) or the fp32 stencil copies
// (mixed), then tau0 = tau (3616)
static void predict_ab(mhd_ctx* c) {
    const long nc = c->ncell, nbi = c->nbi, nbj = c->nbj, nbk = c->nbk;
    if (mixed(c)) {
        AB11<float> s;
        for (int a = 0; a < 11; ++a) {
            s.x[a] = *state_slot(c, a);
            s.xp[a] = c->xp[a];
        }
        // extrapolated state straight into the fp32 stencil copies (sh order: .. bz vf bi bj bk)
        for (int a = 0; a < 8; ++a) s.xh[a] = c->sh[a];
        for (int a = 8; a < 11; ++a) s.xh[a] = c->sh[a + 1];
        const Proj<float> pj = proj<float>(c, c->w32);
        const FaceB fb{c->bi, c->bj, c->bk, c->xp[8], c->xp[9], c->xp[10]};
        const Dims& d = c->d;
        const long npred = nc + (long)d.ny * d.nz + (long)d.nx * d.nz + (long)d.nx * d.ny;
        launch(c, MHD_K_PREV, blocks(npred), dim3(kBlock), k_predict_ab_mixed, c->d, s, fb, bmet(c, false, false),
               (const double*)(c->scal + SC_TAU), c->sh[8], pj, nc, nbi, nbj, nbk);
        launch(c, MHD_K_PREV, dim3(1), dim3(64), k_tau0_from_tau, c->scal + SC_TAU);
    } else {
        AB11<double> s;
        for (int a = 0; a < 11; ++a) {
            s.x[a] = *state_slot(c, a);
            s.xp[a] = c->xp[a];
            s.xh[a] = c->xh[a];
        }
        const Proj<double> pj = proj<double>(c, c->wh64);
        launch(c, MHD_K_PREV, blocks(c->nmax), dim3(kBlock), k_predict_ab<double>, s,
               (const double*)(c->scal + SC_TAU), c->vfh, pj, nc, nbi, nbj, nbk);
        launch(c, MHD_K_PREV, dim3(1), dim3(64), k_tau0_from_tau, c->scal + SC_TAU);
    }
}

static void step_ab(mhd_ctx* c) {
    do_dt(c, true);
    predict_ab(c);
    const bool fused = can_fuse(c);
    if (fused) stage_fused(c, c->xn);
    else do_stage<0, 0>(c, h_src(c), c->xn);
    // rotate: previous <- current <- next <- previous
    for (int a = 0; a < 11; ++a) {
        double* cur = *state_slot(c, a);
        *state_slot(c, a) = c->xn[a];
        c->xn[a] = c->xp[a];
        c->xp[a] = cur;
    }
    do_fill(c);
    if (fuse_dt(c)) {  // the update left the CFL partials of the new state
        c->dt_fresh = true;
        c->ndt_valid = fused ? fused_blocks(c->d) : blocks(c->ninterior).x;
        c->cellb_stale = true;
    } else {
        end_of_stage(c, true);
    }
}

int mhd_step(mhd_ctx* c, int nsteps, double* tau, double* tau_cfl) {
    if (!c || nsteps < 0 || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    if (c->p.stepop != 1 && tau) c->tau0_host = *tau;
    if (nsteps > 0) c->pred_valid = false;
    for (int n = 0; n < nsteps; ++n) {
        if (c->p.stepop == 1) step_ab(c);
        else if (c->p.stepop == 2) step_rk<2>(c);
        else step_rk<3>(c);
    }
    double h[3] = {0., 0., 0.};
    if (hipMemcpyAsync(h, c->scal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    int rc = sync_status(c);
    if (tau) *tau = h[SC_TAU];
    if (tau_cfl) *tau_cfl = h[SC_DTCFL];
    return rc;
}

// ---------------------------------------------------------------------------
// per-routine drop-ins of the time stepper (the calls step_ab / step_rk2 /
// step_rk3 make between get_dt and the ghost fill)
// ---------------------------------------------------------------------------
int mhd_set_tau(mhd_ctx* c, double tau, double tau0) {
    if (!c || !(tau > 0.) || !(tau0 > 0.) || !std::isfinite(tau) || !std::isfinite(tau0)) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    launch(c, MHD_K_PREV, dim3(1), dim3(64), k_set2, c->scal + SC_TAU, tau, tau0);
    c->tau0_host = tau0;
    return sync_status(c);
}

int mhd_get_tau(mhd_ctx* c, double* tau, double* tau0) {
    if (!c) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    double h[2] = {0., 0.};
    if (hipMemcpyAsync(h, c->scal + SC_TAU, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    const int rc = sync_status(c);
    if (tau) *tau = h[0];
    if (tau0) *tau0 = h[1];
    return rc;
}

int mhd_predict_ab(mhd_ctx* c) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    if (c->p.stepop != 1) return MHD_ERR_UNSUPPORTED;
    hipSetDevice(c->p.device);
    if (!mixed(c)) ensure_cellb(c);   // the binary64 predictor extrapolates cell B too
    predict_ab(c);
    // advance_previous (3794-3819, Adams-Bashforth flavour): x_p = x
    for (int a = 0; a < 11; ++a) {
        const long n = a < 8 ? c->ncell : (a == 8 ? c->nbi : (a == 9 ? c->nbj : c->nbk));
        if (hipMemcpyAsync(c->xp[a], *state_slot(c, a), n * sizeof(double), hipMemcpyDeviceToDevice, c->s) !=
            hipSuccess)
            return MHD_ERR_HIP;
    }
    const int rc = sync_status(c);
    c->pred_valid = rc == MHD_OK;
    return rc;
}

// one stage's update in place, in the reference's serial order
template <int NST, int ST>
static void advance_serial(mhd_ctx* c) {
    double* self[11];
    for (int a = 0; a < 11; ++a) self[a] = *state_slot(c, a);
    if (mixed(c)) {
        do_cells_t<float, NST, ST>(c, self, false);
        do_ct_t<float, NST, ST>(c, self);
    } else {
        do_cells_t<double, NST, ST>(c, self, false);
        do_ct_t<double, NST, ST>(c, self);
    }
}

int mhd_advance(mhd_ctx* c, int stage) {
    if (!c || !c->have_geometry) return MHD_ERR_PARAM;
    const int nst = c->p.stepop == 1 ? 0 : c->p.stepop;
    if (nst == 0 ? stage != 0 : (stage < 1 || stage > nst)) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    if (stage == 1) do_prev_rk(c);   // advance_previous: the state stage 1 starts from
    if (nst == 0) advance_serial<0, 0>(c);
    else if (nst == 2 && stage == 1) advance_serial<2, 1>(c);
    else if (nst == 2) advance_serial<2, 2>(c);
    else if (stage == 1) advance_serial<3, 1>(c);
    else if (stage == 2) advance_serial<3, 2>(c);
    else advance_serial<3, 3>(c);
    // the reference's cell B and CFL partials follow at get_bxyz
    c->dt_fresh = false;
    c->cellb_stale = true;
    c->pred_valid = false;
    return sync_status(c);
}

int mhd_store(mhd_ctx* c, int which, const double* host_in) {
    if (!c || !host_in) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    const bool f32 = mixed(c);
    void* dst;
    long n;
    if (which >= MHD_FETCH_FLUX && which < MHD_FETCH_FLUX + 24) {
        const int dq = which / 8, fq = which % 8;
        n = c->nf[dq];
        dst = (char*)c->flux[dq] + (size_t)fq * n * (f32 ? 4 : 8);
    } else if (which >= MHD_FETCH_EI && which <= MHD_FETCH_EK) {
        n = c->ne[which - MHD_FETCH_EI];
        dst = c->emf[which - MHD_FETCH_EI];
    } else {
        return MHD_ERR_PARAM;
    }
    if (f32) {
        std::vector<float> tmp(host_in, host_in + n);
        if (hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess)
            return MHD_ERR_HIP;
        return sync_status(c);
    }
    if (hipMemcpyAsync(dst, host_in, n * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    return sync_status(c);
}

// ---------------------------------------------------------------------------
// batched ensemble: one graph whose branches are the members' RK steps
// ---------------------------------------------------------------------------
struct mhd_ens {
    std::vector<mhd_ctx*> m;
    int device = 0;
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

void mhd_ensemble_destroy(mhd_ens* e) {
    if (!e) return;
    hipSetDevice(e->device);
    if (e->root) hipStreamSynchronize(e->root);
    if (e->exec) hipGraphExecDestroy(e->exec);
    if (e->graph) hipGraphDestroy(e->graph);
    for (hipEvent_t j : e->join)
        if (j) hipEventDestroy(j);
    if (e->fork) hipEventDestroy(e->fork);
    if (e->root) hipStreamDestroy(e->root);
    delete e;
}

int mhd_ensemble_create(mhd_ens** out, mhd_ctx* const* members, int n, const double* tau) {
    if (!out || !members || !tau || n < 1) return MHD_ERR_PARAM;
    *out = nullptr;
    const int dev = members[0] ? members[0]->p.device : 0;
    for (int a = 0; a < n; ++a) {
        mhd_ctx* c = members[a];
        if (!c || !c->have_geometry || c->p.device != dev) return MHD_ERR_PARAM;
        for (int b = 0; b < a; ++b)
            if (members[b] == c) return MHD_ERR_PARAM;
        if (c->p.stepop == 1 || c->p.surrogate) return MHD_ERR_UNSUPPORTED;
    }
    if (hipSetDevice(dev) != hipSuccess) return MHD_ERR_HIP;
    mhd_ens* e = new mhd_ens();
    e->m.assign(members, members + n);
    e->device = dev;
    e->join.assign(n, nullptr);
    bool ok = hipStreamCreateWithFlags(&e->root, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
    for (int a = 0; ok && a < n; ++a) ok = hipEventCreateWithFlags(&e->join[a], hipEventDisableTiming) == hipSuccess;
    // finish whatever the members have queued, then record one step of each
    // member on its own stream, forked from and joined into the root stream
    for (int a = 0; ok && a < n; ++a) ok = hipStreamSynchronize(e->m[a]->s) == hipSuccess;
    if (!ok) {
        mhd_ensemble_destroy(e);
        return MHD_ERR_HIP;
    }
    std::vector<bool> prof(n), ovl(n);
    for (int a = 0; a < n; ++a) {
        prof[a] = e->m[a]->prof;
        ovl[a] = e->m[a]->overlap;
        e->m[a]->prof = false;      // no event records inside the graph
        e->m[a]->overlap = false;   // the members' branches already run side by side
        e->m[a]->tau0_host = tau[a];
    }
    ok = hipStreamBeginCapture(e->root, hipStreamCaptureModeThreadLocal) == hipSuccess &&
         hipEventRecord(e->fork, e->root) == hipSuccess;
    for (int a = 0; ok && a < n; ++a) {
        mhd_ctx* c = e->m[a];
        ok = hipStreamWaitEvent(c->s, e->fork, 0) == hipSuccess;
        if (!ok) break;
        if (c->p.stepop == 2) step_rk<2>(c);
        else step_rk<3>(c);
        ok = c->launch_error == hipSuccess && hipEventRecord(e->join[a], c->s) == hipSuccess &&
             hipStreamWaitEvent(e->root, e->join[a], 0) == hipSuccess;
    }
    hipGraph_t g = nullptr;
    const bool ended = hipStreamEndCapture(e->root, &g) == hipSuccess;
    e->graph = g;
    for (int a = 0; a < n; ++a) {
        e->m[a]->prof = prof[a];
        e->m[a]->overlap = ovl[a];
        e->m[a]->launch_error = hipSuccess;
    }
    ok = ok && ended && g && hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) {
        mhd_ensemble_destroy(e);
        return MHD_ERR_HIP;
    }
    *out = e;
    return MHD_OK;
}

int mhd_ensemble_step(mhd_ens* e, int nsteps) {
    if (!e || nsteps < 0) return MHD_ERR_PARAM;
    hipSetDevice(e->device);
    // members' own queued work first, then the graph on the root stream
    for (mhd_ctx* c : e->m)
        if (hipStreamSynchronize(c->s) != hipSuccess) return MHD_ERR_HIP;
    for (int s = 0; s < nsteps; ++s)
        if (hipGraphLaunch(e->exec, e->root) != hipSuccess) return MHD_ERR_HIP;
    return hipStreamSynchronize(e->root) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

int mhd_fetch(mhd_ctx* c, int which, double* out) {
    if (!c || !out) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    const bool f32 = c->p.precision == MHD_MIXED;
    const void* src;
    long n;
    bool typed = true;
    if (which >= MHD_FETCH_FLUX && which < MHD_FETCH_FLUX + 24) {
        const int dq = which / 8, fq = which % 8;
        n = c->nf[dq];
        src = (const char*)c->flux[dq] + (size_t)fq * n * (f32 ? 4 : 8);
    } else if (which >= MHD_FETCH_EI && which <= MHD_FETCH_EK) {
        n = c->ne[which - MHD_FETCH_EI];
        src = c->emf[which - MHD_FETCH_EI];
    } else if (which == MHD_FETCH_VF) {
        n = c->ncell;
        src = c->vf;
        typed = false;
    } else {
        return MHD_ERR_PARAM;
    }
    if (f32 && typed) {
        std::vector<float> tmp(n);
        if (hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
            return MHD_ERR_HIP;
        int rc = sync_status(c);
        for (long q = 0; q < n; ++q) out[q] = tmp[q];
        return rc;
    }
    if (hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess) return MHD_ERR_HIP;
    return sync_status(c);
}

// arguments of reconstruct_3d: every stencil tap must stay inside f
static int recon_args(const mhd_ctx* c, int dir, const double* f, int fn1, int fn2, int fn3, int lnx, int lny,
                      int lnz, const double* fl, const double* fr) {
    if (!c || !f || !fl || !fr || dir < 1 || dir > 6 || fn1 <= 0 || fn2 <= 0 || fn3 <= 0 || lnx <= 0 ||
        lny <= 0 || lnz <= 0)
        return MHD_ERR_PARAM;
    const int ng = c->p.ng;
    const int lo = c->p.reconop == 1 ? -4 : (c->p.reconop == 2 ? -2 : -3);
    const int hi = c->p.reconop == 1 ? 3 : (c->p.reconop == 2 ? 1 : 2);
    const int ax = dir <= 3 ? dir : dir - 3;
    const int l[3] = {lnx, lny, lnz}, fn[3] = {fn1, fn2, fn3};
    for (int a = 0; a < 3; ++a) {
        const int off = (dir <= 3 || a == ax - 1) ? ng : 0;
        const int mn = 1 + off + (a == ax - 1 ? lo : 0), mx = l[a] + off + (a == ax - 1 ? hi : 0);
        if (mn < 1 || mx > fn[a]) return MHD_ERR_PARAM;
    }
    return MHD_OK;
}

int mhd_reconstruct(mhd_ctx* c, int dir, const double* f, int fn1, int fn2, int fn3, int lnx, int lny, int lnz,
                    double* fl, double* fr) {
    if (int rc = recon_args(c, dir, f, fn1, fn2, fn3, lnx, lny, lnz, fl, fr)) return rc;
    hipSetDevice(c->p.device);
    const size_t nf = (size_t)fn1 * fn2 * fn3, nl = (size_t)lnx * lny * lnz;
    double *df = nullptr, *dl = nullptr, *dr = nullptr;
    if (hipMalloc(&df, nf * sizeof(double)) != hipSuccess) return MHD_ERR_HIP;
    if (hipMalloc(&dl, nl * sizeof(double)) != hipSuccess || hipMalloc(&dr, nl * sizeof(double)) != hipSuccess) {
        hipFree(df);
        if (dl) hipFree(dl);
        return MHD_ERR_HIP;
    }
    int rc = MHD_OK;
    if (hipMemcpyAsync(df, f, nf * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess) rc = MHD_ERR_HIP;
    if (rc == MHD_OK) {
        launch_recon(c, dir, df, fn1, fn2, lnx, lny, lnz, dl, dr);
        if (hipMemcpyAsync(fl, dl, nl * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess ||
            hipMemcpyAsync(fr, dr, nl * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess)
            rc = MHD_ERR_HIP;
        int r2 = sync_status(c);
        if (rc == MHD_OK) rc = r2;
    }
    hipFree(df);
    hipFree(dl);
    hipFree(dr);
    return rc;
}

// reconstruct_3d staged through this GPU's scratch: f, fl and fr take its
// first nf, next nl and last nl doubles (each at a 256-byte boundary)
int mhd_reconstruct_noalloc(mhd_ctx* c, int dir, const double* f, int fn1, int fn2, int fn3, int lnx, int lny,
                            int lnz, double* fl, double* fr) {
    if (int rc = recon_args(c, dir, f, fn1, fn2, fn3, lnx, lny, lnz, fl, fr)) return rc;
    const long nf = (long)fn1 * fn2 * fn3, nl = (long)lnx * lny * lnz;
    const long nf32 = (nf + 31) & ~31L, nl32 = (nl + 31) & ~31L;
    if (nf32 + 2 * nl32 > kScratchDoubles) return MHD_ERR_PARAM;
    if (hipSetDevice(c->p.device) != hipSuccess) return MHD_ERR_HIP;
    double* base = scratch_on(c->p.device);
    if (!base) return MHD_ERR_HIP;
    double *df = base, *dl = base + nf32, *dr = base + nf32 + nl32;
    if (hipMemcpyAsync(df, f, nf * sizeof(double), hipMemcpyHostToDevice, c->s) != hipSuccess) return MHD_ERR_HIP;
    launch_recon(c, dir, df, fn1, fn2, lnx, lny, lnz, dl, dr);
    int rc = MHD_OK;
    if (hipMemcpyAsync(fl, dl, nl * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess ||
        hipMemcpyAsync(fr, dr, nl * sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        rc = MHD_ERR_HIP;
    const int r2 = sync_status(c);
    return rc == MHD_OK ? r2 : rc;
}

int mhd_v_face(mhd_ctx* c, long n, const double* lam, const double* v, double* v0p, double* v0n, double* v1p,
               double* v1n) {
    if (!c || n < 0 || (n > 0 && (!lam || !v || !v0p || !v0n || !v1p || !v1n))) return MHD_ERR_PARAM;
    if (n == 0) return MHD_OK;
    hipSetDevice(c->p.device);
    double* buf = nullptr;
    const size_t b = (size_t)n * sizeof(double);
    if (hipMalloc(&buf, 6 * b) != hipSuccess) return MHD_ERR_HIP;
    double* d[6] = {buf, buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n, buf + 5 * n};
    int rc = MHD_OK;
    if (hipMemcpyAsync(d[0], lam, b, hipMemcpyHostToDevice, c->s) != hipSuccess ||
        hipMemcpyAsync(d[1], v, b, hipMemcpyHostToDevice, c->s) != hipSuccess)
        rc = MHD_ERR_HIP;
    if (rc == MHD_OK) {
        if (mixed(c)) launch_vface_fp32(c, n, d[0], d[1], d[2], d[3], d[4], d[5]);
        else launch_vface_fp64(c, n, d[0], d[1], d[2], d[3], d[4], d[5]);
        double* out[4] = {v0p, v0n, v1p, v1n};
        for (int q = 0; q < 4 && rc == MHD_OK; ++q)
            if (hipMemcpyAsync(out[q], d[2 + q], b, hipMemcpyDeviceToHost, c->s) != hipSuccess) rc = MHD_ERR_HIP;
        const int r2 = sync_status(c);
        if (rc == MHD_OK) rc = r2;
    }
    hipFree(buf);
    return rc;
}

int mhd_divb_max(mhd_ctx* c, double* divb) {
    if (!c || !divb || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    const dim3 g = blocks(c->ninterior);
    auto k = c->met.mode == 2 ? k_divb_partial<true> : k_divb_partial<false>;
    // partials of its own: the CFL partials a step may have left stay valid
    launch(c, MHD_K_DT, g, dim3(kBlock), k, c->d, CFaces{c->bi, c->bj, c->bk}, sdx(c), sdy(c), sdz(c), gmet(c),
           c->divpart, c->ninterior);
    launch(c, MHD_K_DT, dim3(1), dim3(1024), k_dt_final, (const double*)c->divpart, (long)g.x, c->scal + SC_DIVB,
           (double*)nullptr);
    double h = 0.;
    if (hipMemcpyAsync(&h, c->scal + SC_DIVB, sizeof(double), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    const int rc = sync_status(c);
    *divb = -h;
    return rc;
}

int mhd_check_state(mhd_ctx* c, int* flags) {
    if (!c) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    launch(c, MHD_K_DT, dim3(1), dim3(64), k_zero_int<0>, c->flags);
    launch(c, MHD_K_DT, blocks(c->ncell), dim3(kBlock), k_check, c->d, (const double*)c->rho, (const double*)c->pr,
           c->flags, c->ncell);
    int h = 0;
    if (hipMemcpyAsync(&h, c->flags, sizeof(int), hipMemcpyDeviceToHost, c->s) != hipSuccess) return MHD_ERR_HIP;
    int rc = sync_status(c);
    if (flags) *flags = h;
    if (rc != MHD_OK) return rc;
    return h ? MHD_ERR_STATE : MHD_OK;
}

int mhd_synchronize(mhd_ctx* c) {
    if (!c) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    return sync_status(c);
}

void* mhd_device_pointer(mhd_ctx* c, int which) {
    if (!c) return nullptr;
    if (which >= 5 && which <= 7 && c->cellb_stale) {  // cell B: materialise it first
        hipSetDevice(c->p.device);
        ensure_cellb(c);
        if (hipStreamSynchronize(c->s) != hipSuccess) return nullptr;
    }
    void* t[11] = {c->rho, c->vx, c->vy, c->vz, c->pr, c->bx, c->by, c->bz, c->bi, c->bj, c->bk};
    return which >= 0 && which < 11 ? t[which] : nullptr;
}

int mhd_profile(mhd_ctx* c, int enable) {
    if (!c) return MHD_ERR_PARAM;
    c->prof = enable != 0;
    c->prof_mask = ~0u;
    return MHD_OK;
}

int mhd_set_stage_overlap(mhd_ctx* c, int on) {
    if (!c) return MHD_ERR_PARAM;
    c->overlap = on != 0;
    return MHD_OK;
}

int mhd_set_fusion(mhd_ctx* c, int on) {
    if (!c) return MHD_ERR_PARAM;
    c->fused = on != 0;
    return MHD_OK;
}

int mhd_profile_classes(mhd_ctx* c, unsigned mask) {
    if (!c) return MHD_ERR_PARAM;
    c->prof = mask != 0;
    c->prof_mask = mask;
    return MHD_OK;
}

int mhd_profile_read(mhd_ctx* c, int kclass, int* launches, double* ms) {
    if (!c || kclass < 0 || kclass >= MHD_K_NCLASS) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    if (hipStreamSynchronize(c->s) != hipSuccess) return MHD_ERR_HIP;
    for (auto& r : c->evs) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) return MHD_ERR_HIP;
        c->prof_ms[r.kclass] += t;
        c->prof_n[r.kclass] += 1;
        c->evpool.push_back(r.a);
        c->evpool.push_back(r.b);
    }
    c->evs.clear();
    if (launches) *launches = c->prof_n[kclass];
    if (ms) *ms = c->prof_ms[kclass];
    return MHD_OK;
}
