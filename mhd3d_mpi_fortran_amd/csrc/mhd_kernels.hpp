// HIP kernels of the ideal-MHD step (gfx950).
//
// State layout in HBM is the reference's (mhd3d3.f90:251-317): column-major,
// x fastest, ghost-inclusive cell arrays (nx,ny,nz), face arrays
// bi(nx+1,ny,nz) bj(nx,ny+1,nz) bk(nx,ny,nz+1); fluxes and edge EMFs are
// stored without ghosts, one contiguous block per direction.
//
// Every kernel maps one thread to one output point (face, edge, cell) along a
// flattened x-fastest index, so neighbouring lanes of a wavefront touch
// neighbouring addresses for every stencil tap; the stencil reuse across
// lanes and rows is served by L1/L2.  R is the arithmetic type of the
// stencil/flux/EMF work, TF the storage type of fluxes and EMFs; the
// conserved-state accumulation, the stage combine and the constrained-
// transport face update are always binary64.
#pragma once
#include "mhd_numerics.hpp"

namespace mhd {

struct Dims {
    int nx0, ny0, nz0, ng, nx, ny, nz;
};

constexpr int kBlock = 256;

// flat index of 1-based (i,j,k) in an array with leading extents n1, n2
__device__ __forceinline__ ptrdiff_t ix(int n1, int n2, int i, int j, int k) {
    return (ptrdiff_t)(i - 1) + (ptrdiff_t)n1 * ((ptrdiff_t)(j - 1) + (ptrdiff_t)n2 * (k - 1));
}

// 1-based (i,j,k) of flat index t over extents (a, b, *)
// flat index -> 1-based (i, j, k) of an (a, b, *) box; 32-bit unsigned
// arithmetic (mhd_create keeps every index space below 2^31)
__device__ __forceinline__ void unflat(long t, int a, int b, int& i, int& j, int& k) {
    const unsigned u = (unsigned)t, ua = (unsigned)a, ub = (unsigned)b;
    const unsigned r = u / ua;
    i = (int)(u - r * ua) + 1;
    const unsigned s = r / ub;
    j = (int)(r - s * ub) + 1;
    k = (int)s + 1;
}

__device__ __forceinline__ long gtid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }

// XCD-aware block order.  The dispatcher hands consecutive workgroups to the
// 8 XCDs in turn, and each XCD has its own L2: a stencil whose neighbouring
// rows sit in neighbouring workgroups would fetch every row into several
// L2s.  This bijection of [0, nblocks) splits every run of 8*kXcdRun logical
// blocks into 8 sub-runs of kXcdRun, one per XCD: row neighbours share an
// L2, while all XCDs still work inside one window (plane neighbours meet
// in the shared last-level cache).
// RUN is in workgroups of 256 points; a kernel whose workgroups cover G x 256
// points passes 16 / G, so a run always spans 4096 points.
constexpr unsigned kXcd = 8, kXcdRun = 16;
template <unsigned RUN = kXcdRun>
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned b = blockIdx.x, n = gridDim.x;
    const unsigned base = b / (kXcd * RUN) * (kXcd * RUN), r = b - base;
    const unsigned m = n - base < kXcd * RUN ? n - base : kXcd * RUN;
    const unsigned n8 = m / kXcd, rem = m % kXcd, x = r % kXcd, q = r / kXcd;
    return base + x * n8 + (x < rem ? x : rem) + q;
}
__device__ __forceinline__ long gtid_xcd() { return (long)xcd_block() * blockDim.x + threadIdx.x; }

// Face tiles of the strided sweeps (D = 2, 3).  A workgroup's 256 points are
// 64 consecutive faces along x (one wavefront) in 4 consecutive planes along
// the sweep, so the four wavefronts' tap rows overlap -- a W-tap window
// touches W+3 rows per array instead of 4W -- and repeat in the CU's L1
// rather than in L2.  Tiles run x-fastest, then along the sweep, then across
// it, so a run of consecutive tiles on one XCD also shares its rows in L2.
// (ii, jj, kk) = the direction's face counts.
constexpr int kTileX = 64, kTileN = kBlock / kTileX;
template <int D>
__host__ __device__ __forceinline__ unsigned face_tiles(int ii, int jj, int kk) {
    const unsigned nn = D == 2 ? jj : kk, no = D == 2 ? kk : jj;
    return (unsigned)((ii + kTileX - 1) / kTileX) * ((nn + kTileN - 1) / kTileN) * no;
}
// 1-based face (i, j, k) of the calling thread in tile `tile` (the
// convention of unflat and ix); false past the edges
template <int D>
__device__ __forceinline__ bool tile_face(unsigned tile, int ii, int jj, int kk, int& i, int& j, int& k) {
    const unsigned nn = D == 2 ? jj : kk, no = D == 2 ? kk : jj;
    const unsigned ntx = (ii + kTileX - 1) / kTileX, ntn = (nn + kTileN - 1) / kTileN;
    const unsigned tx = tile % ntx, r = tile / ntx, tn = r % ntn, o = r / ntn;
    const unsigned x = tx * kTileX + (threadIdx.x & (kTileX - 1)), s = tn * kTileN + threadIdx.x / kTileX;
    i = (int)x + 1;
    j = (int)(D == 2 ? s : o) + 1;
    k = (int)(D == 2 ? o : s) + 1;
    return x < (unsigned)ii && s < nn && o < no;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// minimum over the workgroup (every thread calls it; two barriers)
__device__ __forceinline__ double block_min(double v) {
    __shared__ double red[kBlock / 64];
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = red[q] < m ? red[q] : m;
        red[0] = m;
    }
    __syncthreads();
    return red[0];
}

// the CFL step of one cell (get_dt, mhd3d3.f90:1608-1628)
__device__ __forceinline__ double cfl_dt(double gam, double cfl, double u, double v, double w, double x, double y,
                                         double z, double r, double pp, double dxi, double dyj, double dzk) {
    const double vv = sqrt(u * u + v * v + w * w);
    const double btot = sqrt(x * x + y * y + z * z);
    const double valf = btot / sqrt(r);
    const double vs = sqrt(gam * pp / r);
    const double vcfl = vv + sqrt(valf * valf + vs * vs);
    return cfl / (vcfl / dxi + vcfl / dyj + vcfl / dzk);
}

template <typename T>
struct Prim {  // cell-centred state read by the flux and EMF kernels
    const T *rho, *vx, *vy, *vz, *p, *bx, *by, *bz;
};

// ------------------------------------------------------------------------
// face flux and magnetic stress: get_flux_stress (mhd3d3.f90:2385-2560)
// F holds 8 arrays of nface each: rho, rvx, rvy, rvz, eng, sx, sy, sz
// ------------------------------------------------------------------------

// Gaussian-split flux of one face from its left/right states
// l, r = (rho, vx, vy, vz, p, bx, by, bz), written at index t of F
template <typename R, int D>
__device__ __forceinline__ void face_flux_vals(const Coef<R>& c, const R* l, const R* r, R* out) {
    const R rl = l[0], ul = l[1], vl = l[2], wl = l[3], pl = l[4], bxl = l[5], byl = l[6], bzl = l[7];
    const R rr = r[0], ur = r[1], vr = r[2], wr = r[3], pr = r[4], bxr = r[5], byr = r[6], bzr = r[7];
    R v0, v1;
    // left state, positive half-moments
    const R unl = D == 1 ? ul : (D == 2 ? vl : wl);
    R leng = R(.5) * rl * (sq(ul) + sq(vl) + sq(wl)) + qdiv(pl, c.gm1);
    R lam = qdiv(rl, R(2) * pl);
    vface<R>(lam, unl, +1, c.rspi, v0, v1);
    R fr0 = rl * v1, fmx = rl * ul * v1, fmy = rl * vl * v1, fmz = rl * wl * v1;
    R fe = (leng + R(.5) * pl) * v1;
    if (D == 1) fmx = fmx + pl * v0;
    if (D == 2) fmy = fmy + pl * v0;
    if (D == 3) fmz = fmz + pl * v0;
    fe = fe + R(.5) * pl * unl * v0;
    R bsq = sq(bxl) + sq(byl) + sq(bzl);
    lam = qdiv(rl, R(2) * pl + bsq);
    vface<R>(lam, unl, +1, c.rspi, v0, v1);
    const R bnl = D == 1 ? bxl : (D == 2 ? byl : bzl);
    R sx = (D == 1 ? (R(.5) * bsq - bxl * bxl) : (-bxl * bnl)) * v0;
    R sy = (D == 2 ? (R(.5) * bsq - byl * byl) : (-byl * bnl)) * v0;
    R sz = (D == 3 ? (R(.5) * bsq - bzl * bzl) : (-bzl * bnl)) * v0;

    // right state, negative half-moments
    const R unr = D == 1 ? ur : (D == 2 ? vr : wr);
    leng = R(.5) * rr * (sq(ur) + sq(vr) + sq(wr)) + qdiv(pr, c.gm1);
    lam = qdiv(rr, R(2) * pr);
    vface<R>(lam, unr, -1, c.rspi, v0, v1);
    R gr0 = rr * v1, gmx = rr * ur * v1, gmy = rr * vr * v1, gmz = rr * wr * v1;
    R ge = (leng + R(.5) * pr) * v1;
    if (D == 1) gmx = gmx + pr * v0;
    if (D == 2) gmy = gmy + pr * v0;
    if (D == 3) gmz = gmz + pr * v0;
    ge = ge + R(.5) * pr * unr * v0;
    bsq = sq(bxr) + sq(byr) + sq(bzr);
    lam = qdiv(rr, R(2) * pr + bsq);
    vface<R>(lam, unr, -1, c.rspi, v0, v1);
    const R bnr = D == 1 ? bxr : (D == 2 ? byr : bzr);
    R tx = (D == 1 ? (R(.5) * bsq - bxr * bxr) : (-bxr * bnr)) * v0;
    R ty = (D == 2 ? (R(.5) * bsq - byr * byr) : (-byr * bnr)) * v0;
    R tz = (D == 3 ? (R(.5) * bsq - bzr * bzr) : (-bzr * bnr)) * v0;

    out[0] = fr0 + gr0;
    out[1] = fmx + gmx;
    out[2] = fmy + gmy;
    out[3] = fmz + gmz;
    out[4] = fe + ge;
    out[5] = sx + tx;
    out[6] = sy + ty;
    out[7] = sz + tz;
}

template <typename R, int D, typename TF>
__device__ __forceinline__ void face_flux(const Coef<R>& c, const R* l, const R* r, TF* __restrict__ F, long t,
                                          long nface) {
    R o[8];
    face_flux_vals<R, D>(c, l, r, o);
#pragma unroll
    for (int q = 0; q < 8; ++q) F[q * nface + t] = TF(o[q]);
}

// ------------------------------------------------------------------------
// affine metric (no counterpart in the reference, whose grid is a tensor
// product, mhd3d3.f90:3483-3485): physical x = A * (computational x), the
// computational grid being the tensor-product one.  Face family d has the
// unit normal n_d = C e_d / |C e_d| (C = cofactor matrix of A); the flux
// through it is the Gaussian-split flux in vector form along n_d.  With n_d
// a coordinate axis this is the reference's get_flux_stress term for term.
// ------------------------------------------------------------------------
// general metric (mhd_set_metric_nodes): per-face and per-cell geometry
// from the node coordinates, binary64, device arrays.  Face arrays use the
// ghost-inclusive face layouts of bi, bj, bk; edge arrays the owned-edge
// layouts of ei, ej, ek; cell arrays the ghost-inclusive cell layout.
struct GMet {
    const double* nf[3];   // unit normals of face family d, 3 components (SoA, stride = face count)
    const double* af[3];   // face areas |S|
    const double* iaf[3];  // 1/|S|
    const double* ef[3];   // EMF factor |S1||S2|/V of the local edge basis, per owned edge (E . dl)
    const double* ivol;    // 1/V
    const double* cm;      // cell B = sum_d beta_d(-) cm[d][0] + beta_d(+) cm[d][1]; 18 per cell (SoA)
    const double* nc;      // cell normals of the 3 face families, 9 per cell (SoA)
    const double* ih;      // CFL: 0.5 (|S-| + |S+|) / V per family, 3 per cell (SoA)
    long ncell, nb[3];
};

template <typename R>
struct Normals {
    R n[3][3];  // n[d] = unit normal of the faces normal to computational axis d
};

template <typename R>
__device__ __forceinline__ void face_flux_met_vals(const Coef<R>& c, const R* l, const R* r, const R* n, R* fx) {
    R v0, v1;
    // left state, positive half-moments
    {
        const R rl = l[0], ul = l[1], vl = l[2], wl = l[3], pl = l[4], bxl = l[5], byl = l[6], bzl = l[7];
        const R un = n[0] * ul + n[1] * vl + n[2] * wl;
        const R leng = R(.5) * rl * (sq(ul) + sq(vl) + sq(wl)) + qdiv(pl, c.gm1);
        R lam = qdiv(rl, R(2) * pl);
        vface<R>(lam, un, +1, c.rspi, v0, v1);
        fx[0] = rl * v1;
        fx[1] = rl * ul * v1 + pl * n[0] * v0;
        fx[2] = rl * vl * v1 + pl * n[1] * v0;
        fx[3] = rl * wl * v1 + pl * n[2] * v0;
        fx[4] = (leng + R(.5) * pl) * v1 + R(.5) * pl * un * v0;
        const R bsq = sq(bxl) + sq(byl) + sq(bzl);
        lam = qdiv(rl, R(2) * pl + bsq);
        vface<R>(lam, un, +1, c.rspi, v0, v1);
        const R bn = n[0] * bxl + n[1] * byl + n[2] * bzl;
        fx[5] = (R(.5) * bsq * n[0] - bxl * bn) * v0;
        fx[6] = (R(.5) * bsq * n[1] - byl * bn) * v0;
        fx[7] = (R(.5) * bsq * n[2] - bzl * bn) * v0;
    }
    // right state, negative half-moments
    {
        const R rr = r[0], ur = r[1], vr = r[2], wr = r[3], pr = r[4], bxr = r[5], byr = r[6], bzr = r[7];
        const R un = n[0] * ur + n[1] * vr + n[2] * wr;
        const R leng = R(.5) * rr * (sq(ur) + sq(vr) + sq(wr)) + qdiv(pr, c.gm1);
        R lam = qdiv(rr, R(2) * pr);
        vface<R>(lam, un, -1, c.rspi, v0, v1);
        fx[0] += rr * v1;
        fx[1] += rr * ur * v1 + pr * n[0] * v0;
        fx[2] += rr * vr * v1 + pr * n[1] * v0;
        fx[3] += rr * wr * v1 + pr * n[2] * v0;
        fx[4] += (leng + R(.5) * pr) * v1 + R(.5) * pr * un * v0;
        const R bsq = sq(bxr) + sq(byr) + sq(bzr);
        lam = qdiv(rr, R(2) * pr + bsq);
        vface<R>(lam, un, -1, c.rspi, v0, v1);
        const R bn = n[0] * bxr + n[1] * byr + n[2] * bzr;
        fx[5] += (R(.5) * bsq * n[0] - bxr * bn) * v0;
        fx[6] += (R(.5) * bsq * n[1] - byr * bn) * v0;
        fx[7] += (R(.5) * bsq * n[2] - bzr * bn) * v0;
    }
}

template <typename R, typename TF>
__device__ __forceinline__ void face_flux_met(const Coef<R>& c, const R* l, const R* r, const R* n,
                                              TF* __restrict__ F, long t, long nface) {
    R fx[8];
    face_flux_met_vals<R>(c, l, r, n, fx);
#pragma unroll
    for (int q = 0; q < 8; ++q) F[q * nface + t] = TF(fx[q]);
}

// normal of face (i, j, k) (owned, 1-based) of family D from a general
// metric's per-face array (ghost-inclusive face layout)
template <typename R, int D>
__device__ __forceinline__ void gen_normal(const Dims& d, const GMet& g, int i, int j, int k, R* n) {
    const ptrdiff_t f = D == 1   ? ix(d.nx + 1, d.ny, d.ng + i, d.ng + j, d.ng + k)
                        : D == 2 ? ix(d.nx, d.ny + 1, d.ng + i, d.ng + j, d.ng + k)
                                 : ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k);
    const double* a = g.nf[D - 1];
    const long nb = g.nb[D - 1];
    n[0] = R(a[f]);
    n[1] = R(a[nb + f]);
    n[2] = R(a[2 * nb + f]);
}

// fp32 strided sweeps (D = 2, 3): a workgroup takes kFluxRun consecutive face
// tiles along the sweep of one x-column run (a tile: 64 x-faces by kTileN
// face planes, one wavefront per plane).  Tile tn needs sweep rows
// 4 tn .. 4 tn + kTileN + n - 2 (row rho holds cell ng + 1 + lo + rho), so
// consecutive tiles share n - 1 rows: the rows sit in an LDS ring of 16 (row
// rho in slot rho mod 16), the first tile loads all of its rows and every
// later one only its kTileN new ones, each row one coalesced load.  Every
// thread of the workgroup calls it, tile after tile (two barriers); columns
// and rows past the block's edges are clamped onto it.  At 256^3 (fp32 PDM,
// sheared) a march of 8 tiles takes the step from 3.10 ms (one tile per
// workgroup, 11 rows loaded per 4 planes) to 2.97 ms; 4 and 16 tiles, and a
// ring of 12 (six workgroups per CU instead of five), measured slower.  A
// march needs kFluxRun times fewer workgroups, so blocks whose sweep would
// then leave fewer than kFluxRunBlocks workgroups take one tile each
// (RUN = 1, the rows in LDS as they come) instead: the shell of 96 x 64 x 128
// and the ensemble's boxes lost 5-25 % with runs of 8.
constexpr int kFluxRun = 8, kFluxRing = 16, kFluxRunBlocks = 4096;
template <int D>
__host__ __device__ __forceinline__ unsigned face_runs(int ii, int jj, int kk, int run) {
    const unsigned nn = D == 2 ? jj : kk, no = D == 2 ? kk : jj;
    const unsigned ntn = (nn + kTileN - 1) / kTileN;
    return (unsigned)((ii + kTileX - 1) / kTileX) * ((ntn + run - 1) / run) * no;
}
template <typename R, int M, int D, typename TI, int RUN>
__device__ __forceinline__ void ring_recon(const Dims& d, const Prim<TI>& q, const Coef<R>& c, unsigned tx,
                                           unsigned tn, unsigned o, int ii, unsigned no, bool first, R* l, R* r) {
    using S = Stencil<R, M>;
    constexpr int NR = kTileN + S::n - 1;
    constexpr int RING = RUN == 1 ? NR : kFluxRing;
    static_assert(NR <= kFluxRing, "the ring holds one tile's rows");
    __shared__ TI win[8][RING][kTileX];
    // LDS slot of the tile's row rho
    auto slot = [&](int rho) { return RUN == 1 ? rho : ((int)(tn * kTileN) + rho) & (kFluxRing - 1); };
    const TI* f[8] = {q.rho, q.vx, q.vy, q.vz, q.p, q.bx, q.by, q.bz};
    const int lane = threadIdx.x & (kTileX - 1), w = threadIdx.x / kTileX;
    const int cx = d.ng + 1 + min((int)(tx * kTileX) + lane, ii - 1);
    const int co = d.ng + 1 + min((int)o, (int)no - 1);
    const int cmax = D == 2 ? d.ny : d.nz;
    const int r0 = (int)(tn * kTileN);
    if (RUN > 1) __syncthreads();  // the previous tile's reads are done
#pragma unroll
    for (int rho = first ? w : NR - kTileN + w; rho < NR; rho += kTileN) {
        const int cs = max(1, min(d.ng + 1 + S::lo + r0 + rho, cmax));
        const unsigned at = (unsigned)(D == 2 ? ix(d.nx, d.ny, cx, cs, co) : ix(d.nx, d.ny, cx, co, cs));
#pragma unroll
        for (int v = 0; v < 8; ++v) win[v][slot(rho)][lane] = f[v][at];
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        R t[S::n];
#pragma unroll
        for (int m = 0; m < S::n; ++m) t[m] = R(win[v][slot(w + m)][lane]);
        S::run(t, c, l[v], r[v]);
    }
}

// MET = 0: the reference's axis-aligned flux; 1: along the affine metric's
// normals; 2: along the general metric's per-face normals
// RUN: face tiles per workgroup of the fp32 strided sweeps
template <typename R, int M, int D, typename TI, typename TF, int MET, int RUN = 1>
__global__ void __launch_bounds__(kBlock) k_flux(Dims d, Coef<R> c, Prim<TI> q, Normals<R> nm, GMet gm,
                                                 TF* __restrict__ F, long nface) {
    const int ii = d.nx0 + (D == 1), jj = d.ny0 + (D == 2), kk = d.nz0 + (D == 3);
    const ptrdiff_t st = D == 1 ? 1 : (D == 2 ? (ptrdiff_t)d.nx : (ptrdiff_t)d.nx * d.ny);
    int i, j, k;
    long t;
    R l[8], r[8];
    if constexpr (D == 1) {
        // flat; unit stride: the taps of one array are contiguous and load
        // as wide vectors from one address
        t = gtid_xcd();
        if (t >= nface) return;
        unflat(t, ii, jj, i, j, k);
        const ptrdiff_t base = ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k);
        recon_line<R, M>(q.rho, base, st, c, l[0], r[0]);
        recon_line<R, M>(q.vx, base, st, c, l[1], r[1]);
        recon_line<R, M>(q.vy, base, st, c, l[2], r[2]);
        recon_line<R, M>(q.vz, base, st, c, l[3], r[3]);
        recon_line<R, M>(q.p, base, st, c, l[4], r[4]);
        recon_line<R, M>(q.bx, base, st, c, l[5], r[5]);
        recon_line<R, M>(q.by, base, st, c, l[6], r[6]);
        recon_line<R, M>(q.bz, base, st, c, l[7], r[7]);
    } else if constexpr (sizeof(TI) == 4) {
        // fp32 stencil inputs: RUN tiles along the sweep, their rows
        // through the LDS ring
        const unsigned nn = D == 2 ? jj : kk, no = D == 2 ? kk : jj;
        const unsigned ntx = (ii + kTileX - 1) / kTileX, ntn = (nn + kTileN - 1) / kTileN;
        const unsigned ntc = (ntn + RUN - 1) / RUN;
        const unsigned lb = xcd_block<kXcdRun / RUN>();
        const unsigned tx = lb % ntx, tnc = lb / ntx % ntc, o = lb / ntx / ntc;
        const unsigned x = tx * kTileX + (threadIdx.x & (kTileX - 1));
#pragma unroll 1
        for (int g = 0; g < RUN; ++g) {
            const unsigned tn = tnc * RUN + g, sw = tn * kTileN + threadIdx.x / kTileX;
            ring_recon<R, M, D, TI, RUN>(d, q, c, tx, tn, o, ii, no, g == 0, l, r);
            if (x < (unsigned)ii && sw < nn) {
                i = (int)x + 1;
                j = (int)(D == 2 ? sw : o) + 1;
                k = (int)(D == 2 ? o : sw) + 1;
                t = ix(ii, jj, i, j, k);
                if (MET == 2) {
                    R n[3];
                    gen_normal<R, D>(d, gm, i, j, k, n);
                    face_flux_met<R, TF>(c, l, r, n, F, t, nface);
                } else if (MET == 1) {
                    face_flux_met<R, TF>(c, l, r, nm.n[D - 1], F, t, nface);
                } else {
                    face_flux<R, D, TF>(c, l, r, F, t, nface);
                }
            }
        }
        return;
    } else {
        // binary64 (VALU-bound): taps from the tile's rows in L1
        if (!tile_face<D>(xcd_block(), ii, jj, kk, i, j, k)) return;
        t = ix(ii, jj, i, j, k);
        const Taps<R, M, TI> tp(ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k), st);
        recon_taps<R, M>(q.rho, tp, c, l[0], r[0]);
        recon_taps<R, M>(q.vx, tp, c, l[1], r[1]);
        recon_taps<R, M>(q.vy, tp, c, l[2], r[2]);
        recon_taps<R, M>(q.vz, tp, c, l[3], r[3]);
        recon_taps<R, M>(q.p, tp, c, l[4], r[4]);
        recon_taps<R, M>(q.bx, tp, c, l[5], r[5]);
        recon_taps<R, M>(q.by, tp, c, l[6], r[6]);
        recon_taps<R, M>(q.bz, tp, c, l[7], r[7]);
    }
    if (MET == 2) {
        R n[3];
        gen_normal<R, D>(d, gm, i, j, k, n);
        face_flux_met<R, TF>(c, l, r, n, F, t, nface);
    } else if (MET == 1) {
        face_flux_met<R, TF>(c, l, r, nm.n[D - 1], F, t, nface);
    } else {
        face_flux<R, D, TF>(c, l, r, F, t, nface);
    }
}

// ------------------------------------------------------------------------
// standalone reconstruct_3d (mhd3d3.f90:833-860) for dir 1..6
// ------------------------------------------------------------------------
template <typename R, int M>
__global__ void __launch_bounds__(kBlock) k_recon(int ng, int dir, const double* __restrict__ f, int fn1, int fn2,
                                                 int lnx, int lny, int lnz, Coef<R> c, double* __restrict__ fl,
                                                 double* __restrict__ fr) {
    const long t = gtid();
    if (t >= (long)lnx * lny * lnz) return;
    int i, j, k;
    unflat(t, lnx, lny, i, j, k);
    const int ax = dir <= 3 ? dir : dir - 3;
    const int off = dir <= 3 ? ng : 0;
    int c3[3] = {i + off, j + off, k + off};
    c3[ax - 1] = (ax == 1 ? i : (ax == 2 ? j : k)) + ng;
    const ptrdiff_t base = ix(fn1, fn2, c3[0], c3[1], c3[2]);
    const ptrdiff_t st = ax == 1 ? 1 : (ax == 2 ? (ptrdiff_t)fn1 : (ptrdiff_t)fn1 * fn2);
    R a, b;
    recon_line<R, M>(f, base, st, c, a, b);
    fl[t] = double(a);
    fr[t] = double(b);
}

}  // namespace mhd
