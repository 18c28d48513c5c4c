// Context object and launch plumbing shared by the library's translation
// units (binary64 stencil kernels, fp32 stencil kernels, surrogate).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mhd_gpu.h"
#include "mhd_kernels.hpp"

namespace mhd {
// stencil inputs of one state: cell primitives, cell B, vf, face B and,
// with a metric, the face-normal velocities w_d = v . n_d
template <typename T>
struct Src {
    const T *rho, *vx, *vy, *vz, *p, *bx, *by, *bz, *vf, *bi, *bj, *bk;
    const T* w[3];
};

// affine metric x_phys = A x_comp (mhd_set_metric); on = false is the
// reference's tensor-product grid, for which every factor below is exactly 1.
// mode 2: the general metric of GMet (mhd_set_metric_nodes)
struct Metric {
    bool on = false;
    int mode = 0;  // 0 tensor product, 1 affine, 2 general
    double A[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // row-major
    double n[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};  // unit face normals per computational axis
    double g[3] = {1, 1, 1};    // |C e_d| / det A: flux divergence factor (update, CFL)
    double ct[3] = {1, 1, 1};   // 1 / |C e_d|: constrained-transport factor of face family d
    double ef[3] = {1, 1, 1};   // |C e_a||C e_b| / det A: EMF factor of edges along d
    double M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // A diag(g): face-average B -> cell B
};

// get_eijk_Yee's twelve split fields e<a><s1><s2> of edge direction a (x, y,
// z), s1 the side of the first transverse velocity's moments, s2 the
// second's: e_a = B_b v_b0(s1) v_c1(s2) - B_c v_b1(s1) v_c0(s2), (a, b, c)
// cyclic; index 4*a + YEE_<s1><s2>
enum { YEE_PP = 0, YEE_NP = 1, YEE_PN = 2, YEE_NN = 3 };
}  // namespace mhd

// device scalars: scal[SC_TAU] = tau, scal[SC_TAU0] = tau0, scal[SC_DTCFL] = get_dt result,
// scal[SC_DIVB] = -max |div B| of the last mhd_divb_max
enum { SC_TAU = 0, SC_TAU0 = 1, SC_DTCFL = 2, SC_DIVB = 3 };

struct SurrogateState;

// EMF launches whose long-run grid would leave the chip short of waves
// (fewer than ~8 per SIMD: 256 CUs x 4 SIMDs x 8 waves x 64 lanes) take
// short runs
constexpr long kEmfFillThreads = 256L * 4 * 8 * 64;
// blocks with at least this many owned cells run a stage on two streams by
// default (mhd_set_stage_overlap)
constexpr long kOverlapCells = 1L << 21;

struct mhd_ctx {
    mhd_params p{};
    mhd::Dims d{};
    hipStream_t s = nullptr;
    // second stream: a stage's edge EMF and CT run on it beside the face flux
    // and the cell update (see stage_overlapped).  cur: the stream launch()
    // targets (null = s)
    hipStream_t s2 = nullptr, cur = nullptr;
    hipEvent_t ev_fork = nullptr, ev_emf = nullptr, ev_join = nullptr;
    bool overlap = true;   // false: one stream, kernels in the reference's order
    bool fused = false;    // mixed Adams-Bashforth PDM steps: fused face flux + update (mhd_set_fusion)
    char* arena = nullptr;
    size_t arena_bytes = 0;
    bool have_geometry = false;
    long ncell = 0, nbi = 0, nbj = 0, nbk = 0, ninterior = 0, ndtpart = 0, nmax = 0;
    long nf[3] = {0, 0, 0}, ne[3] = {0, 0, 0};
    // EMF launches with fewer long-run threads than this take short runs
    // (environment MHD_EMF_FILL_THREADS overrides it at mhd_create)
    long emf_fill = 0;
    // state (reference layout)
    double *rho = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr, *pr = nullptr;
    double *bx = nullptr, *by = nullptr, *bz = nullptr, *vf = nullptr;
    double *bi = nullptr, *bj = nullptr, *bk = nullptr;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr;
    // RK advance_previous
    double *rhop = nullptr, *mxp = nullptr, *myp = nullptr, *mzp = nullptr, *engp = nullptr;
    double *bip = nullptr, *bjp = nullptr, *bkp = nullptr;
    // AB: previous and next state (rotated with the current one each step),
    // extrapolated state (binary64 mode; mixed mode extrapolates into sh)
    // order: rho vx vy vz p bx by bz bi bj bk
    double* xp[11] = {};
    double* xn[11] = {};
    double* xh[11] = {};
    double* vfh = nullptr;
    // fp32 stencil inputs (mixed precision): rho vx vy vz p bx by bz vf bi bj bk
    float* sh[12] = {};
    bool dt_fresh = false;   // dtpart holds the CFL partials of the current state
    long ndt_valid = 0;      // ... that many of them
    bool ghost_twins = false;  // every ghost cell repeats an owned cell (periodic axes, equal ghost spacings)
    bool cellb_stale = false;  // bx,by,bz not stored for the current state (mixed AB steps, uploads, advances)
    bool pred_valid = false;   // the extrapolated state of mhd_predict_ab is held (xh / fp32 copies)
    // per-stage products (binary64, or fp32 in mixed precision)
    void* flux[3] = {nullptr, nullptr, nullptr};
    void* emf[3] = {nullptr, nullptr, nullptr};
    double* dtpart = nullptr;
    double* divpart = nullptr;   // per-workgroup partials of the div B monitor (apart from the CFL ones)
    double* scal = nullptr;
    int* flags = nullptr;
    double tau0_host = 0.;
    hipError_t launch_error = hipSuccess;
    // kernel timing
    bool prof = false;
    unsigned prof_mask = ~0u;   // classes timed while prof is set
    struct Rec {
        int kclass;
        hipEvent_t a, b;
    };
    std::vector<Rec> evs;
    std::vector<hipEvent_t> evpool;
    double prof_ms[MHD_K_NCLASS] = {};
    int prof_n[MHD_K_NCLASS] = {};
    // metric (mhd_set_metric): spacing divided by g_d for the update and the
    // CFL, face-normal velocities (binary64 current / extrapolated state, fp32)
    mhd::Metric met;
    std::vector<double> hdx, hdy, hdz;
    double *dxm = nullptr, *dym = nullptr, *dzm = nullptr;
    char* metbuf = nullptr;
    char* gmetbuf = nullptr;
    mhd::GMet gm{};
    double* w64[3] = {};
    double* wh64[3] = {};
    float* w32[3] = {};
    SurrogateState* sur = nullptr;
    float sur_theta = 0.25f;   // smoothness threshold of the learned reconstruction
    bool sur_stats = false;    // count faces / fallbacks
};

inline dim3 blocks(long n) { return dim3((unsigned)((n + mhd::kBlock - 1) / mhd::kBlock)); }

inline hipEvent_t ev_get(mhd_ctx* c) {
    if (!c->evpool.empty()) {
        hipEvent_t e = c->evpool.back();
        c->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <typename... KArgs, typename... Args>
inline void launch(mhd_ctx* c, int kclass, dim3 g, dim3 b, void (*k)(KArgs...), Args... args) {
    if (g.x == 0 || g.y == 0) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = c->prof && ((c->prof_mask >> kclass) & 1u);
    hipStream_t st = c->cur ? c->cur : c->s;
    auto note = [c](hipError_t err) {
        if (err != hipSuccess && c->launch_error == hipSuccess) c->launch_error = err;
    };
    if (timed) {
        e0 = ev_get(c);
        e1 = ev_get(c);
        if (!e0 || !e1) note(hipErrorOutOfMemory);
        else note(hipEventRecord(e0, st));
    }
    hipLaunchKernelGGL(k, g, b, 0, st, static_cast<KArgs>(args)...);
    note(hipGetLastError());
    if (timed && e0 && e1) {
        note(hipEventRecord(e1, st));
        c->evs.push_back({kclass, e0, e1});
    } else {
        if (e0) c->evpool.push_back(e0);
        if (e1) c->evpool.push_back(e1);
    }
}

template <typename R>
inline mhd::Coef<R> make_coef(const mhd_ctx* c) {
    mhd::Coef<R> k;
    const double pi = 3.141592653589793238462643383279502884;  // binary64 of 4*atan(1)
    k.gam = R(c->p.gam);
    k.gm1 = R(c->p.gam - 1.);
    k.cfl = R(c->p.cfl);
    k.pdmb = R(c->p.pdmb);
    k.eps = R(c->p.eps);
    k.alpha = R(c->p.alpha);
    k.beta = R(c->p.beta);
    k.rspi = R(std::sqrt(pi));
    return k;
}

template <typename R>
inline mhd::Normals<R> make_normals(const mhd_ctx* c) {
    mhd::Normals<R> k;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) k.n[a][b] = R(c->met.n[a][b]);
    return k;
}

// per-precision stencil launchers (separate translation units)
void launch_flux_fp64(mhd_ctx* c, mhd::Src<double> s);
void launch_emf_fp64(mhd_ctx* c, mhd::Src<double> s);
void launch_recon(mhd_ctx* c, int dir, const double* f, int fn1, int fn2, int lnx, int lny, int lnz, double* fl,
                  double* fr);
void launch_flux_fp32(mhd_ctx* c, mhd::Src<float> s);
void launch_emf_fp32(mhd_ctx* c, mhd::Src<float> s);
void launch_flux_surrogate(mhd_ctx* c, mhd::Src<float> s);
void launch_vface_fp64(mhd_ctx* c, long n, const double* lam, const double* v, double* v0p, double* v0n, double* v1p,
                       double* v1n);
void launch_vface_fp32(mhd_ctx* c, long n, const double* lam, const double* v, double* v0p, double* v0n, double* v1p,
                       double* v1n);
void surrogate_free(mhd_ctx* c);
// get_eijk_Yee's split (12 ghost-inclusive cell arrays, e[4*a + YEE_*]) and
// quadrant sums (mhd_yee.hpp)
void launch_yee_split(mhd_ctx* c, long n, mhd::Prim<double> q, double* const* e);
void launch_yee_sum(mhd_ctx* c, int o1, int o2, int o3, int ax, int n1, int n2, const double* const* e4, double* out);
