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
__device__ __forceinline__ void unflat(long t, int a, int b, int& i, int& j, int& k) {
    long r = t / a;
    i = (int)(t - r * a) + 1;
    long s = r / b;
    j = (int)(r - s * b) + 1;
    k = (int)s + 1;
}

__device__ __forceinline__ long gtid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }

template <typename T>
struct Prim {  // cell-centred state read by the flux and EMF kernels
    const T *rho, *vx, *vy, *vz, *p, *bx, *by, *bz;
};

// ------------------------------------------------------------------------
// face flux and magnetic stress: get_flux_stress (mhd3d3.f90:2385-2560)
// F holds 8 arrays of nface each: rho, rvx, rvy, rvz, eng, sx, sy, sz
// ------------------------------------------------------------------------
template <typename R, int M, int D, typename TF>
__global__ void __launch_bounds__(kBlock) k_flux(Dims d, Coef<R> c, Prim<double> q, TF* __restrict__ F,
                                                long nface) {
    const long t = gtid();
    if (t >= nface) return;
    const int ii = d.nx0 + (D == 1), jj = d.ny0 + (D == 2);
    int i, j, k;
    unflat(t, ii, jj, i, j, k);
    const ptrdiff_t base = ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k);
    const ptrdiff_t st = D == 1 ? 1 : (D == 2 ? (ptrdiff_t)d.nx : (ptrdiff_t)d.nx * d.ny);
    R rl, rr, ul, ur, vl, vr, wl, wr, pl, pr, bxl, bxr, byl, byr, bzl, bzr;
    recon_line<R, M>(q.rho, base, st, c, rl, rr);
    recon_line<R, M>(q.vx, base, st, c, ul, ur);
    recon_line<R, M>(q.vy, base, st, c, vl, vr);
    recon_line<R, M>(q.vz, base, st, c, wl, wr);
    recon_line<R, M>(q.p, base, st, c, pl, pr);
    recon_line<R, M>(q.bx, base, st, c, bxl, bxr);
    recon_line<R, M>(q.by, base, st, c, byl, byr);
    recon_line<R, M>(q.bz, base, st, c, bzl, bzr);

    R v0, v1;
    // left state, positive half-moments
    const R unl = D == 1 ? ul : (D == 2 ? vl : wl);
    R leng = R(.5) * rl * (sq(ul) + sq(vl) + sq(wl)) + pl / c.gm1;
    R lam = rl / (R(2) * pl);
    vface<R>(lam, unl, +1, c.rspi, v0, v1);
    R fr0 = rl * v1, fmx = rl * ul * v1, fmy = rl * vl * v1, fmz = rl * wl * v1;
    R fe = (leng + R(.5) * pl) * v1;
    if (D == 1) fmx = fmx + pl * v0;
    if (D == 2) fmy = fmy + pl * v0;
    if (D == 3) fmz = fmz + pl * v0;
    fe = fe + R(.5) * pl * unl * v0;
    R bsq = sq(bxl) + sq(byl) + sq(bzl);
    lam = rl / (R(2) * pl + bsq);
    vface<R>(lam, unl, +1, c.rspi, v0, v1);
    const R bnl = D == 1 ? bxl : (D == 2 ? byl : bzl);
    R sx = (D == 1 ? (R(.5) * bsq - bxl * bxl) : (-bxl * bnl)) * v0;
    R sy = (D == 2 ? (R(.5) * bsq - byl * byl) : (-byl * bnl)) * v0;
    R sz = (D == 3 ? (R(.5) * bsq - bzl * bzl) : (-bzl * bnl)) * v0;

    // right state, negative half-moments
    const R unr = D == 1 ? ur : (D == 2 ? vr : wr);
    leng = R(.5) * rr * (sq(ur) + sq(vr) + sq(wr)) + pr / c.gm1;
    lam = rr / (R(2) * pr);
    vface<R>(lam, unr, -1, c.rspi, v0, v1);
    R gr0 = rr * v1, gmx = rr * ur * v1, gmy = rr * vr * v1, gmz = rr * wr * v1;
    R ge = (leng + R(.5) * pr) * v1;
    if (D == 1) gmx = gmx + pr * v0;
    if (D == 2) gmy = gmy + pr * v0;
    if (D == 3) gmz = gmz + pr * v0;
    ge = ge + R(.5) * pr * unr * v0;
    bsq = sq(bxr) + sq(byr) + sq(bzr);
    lam = rr / (R(2) * pr + bsq);
    vface<R>(lam, unr, -1, c.rspi, v0, v1);
    const R bnr = D == 1 ? bxr : (D == 2 ? byr : bzr);
    R tx = (D == 1 ? (R(.5) * bsq - bxr * bxr) : (-bxr * bnr)) * v0;
    R ty = (D == 2 ? (R(.5) * bsq - byr * byr) : (-byr * bnr)) * v0;
    R tz = (D == 3 ? (R(.5) * bsq - bzr * bzr) : (-bzr * bnr)) * v0;

    F[t] = TF(fr0 + gr0);
    F[nface + t] = TF(fmx + gmx);
    F[2 * nface + t] = TF(fmy + gmy);
    F[3 * nface + t] = TF(fmz + gmz);
    F[4 * nface + t] = TF(fe + ge);
    F[5 * nface + t] = TF(sx + tx);
    F[6 * nface + t] = TF(sy + ty);
    F[7 * nface + t] = TF(sz + tz);
}

// ------------------------------------------------------------------------
// corner (edge) velocity: get_vav_center2corner (mhd3d3.f90:1858-2116)
// b = cell index of (I,J,K), the cell above the edge in both transverse
// axes; s1/s2 = strides along the vav axes d1/d2 of the edge direction.
// ------------------------------------------------------------------------
template <typename R, int M>
__device__ __forceinline__ R vav(const double* __restrict__ lv, ptrdiff_t b, ptrdiff_t s1, ptrdiff_t s2,
                                 const Coef<R>& c) {
    auto cv = [&](int o1, int o2) -> R { return R(lv[b + o1 * s1 + o2 * s2]); };
    R v00_0, v10_0, v01_0, v11_0, v00_1, v01_1, v10_1, v11_1;
    if constexpr (M == PDM) {
        // interp2_pdm: along d1 at 8 levels of d2, then along d2 (926-989)
        R g[8];
#pragma unroll
        for (int m2 = 0; m2 < 8; ++m2) {
            R s[8];
#pragma unroll
            for (int m1 = 0; m1 < 8; ++m1) s[m1] = cv(m1 - 4, m2 - 4);
            g[m2] = interp8(s);
        }
        const R vi = interp8(g);
        // one-sided PDM2 estimates (1899-1912)
        pdm2(cv(-2, -1), cv(-1, -1), cv(0, -1), cv(1, -1), vi, c.pdmb, v00_0, v10_0);
        pdm2(cv(-1, -2), cv(-1, -1), cv(-1, 0), cv(-1, 1), vi, c.pdmb, v00_1, v01_1);
        pdm2(cv(0, -2), cv(0, -1), cv(0, 0), cv(0, 1), vi, c.pdmb, v10_1, v11_1);
        pdm2(cv(-2, 0), cv(-1, 0), cv(0, 0), cv(1, 0), vi, c.pdmb, v01_0, v11_0);
    } else if constexpr (M == TVD) {
        // interp2_tvd (1037-1067); TVD ignores the shifts (1187-1238)
        const R vi = (cv(-1, -1) + cv(0, -1) + cv(0, 0) + cv(-1, 0)) / R(4);
        tvd(cv(-2, 0), cv(-1, 0), cv(0, 0), cv(1, 0), vi, c.eps, v00_0, v10_0);
        tvd(cv(0, -2), cv(0, -1), cv(0, 0), cv(0, 1), vi, c.eps, v00_1, v01_1);
        v10_1 = v00_1;
        v11_1 = v01_1;
        v01_0 = v00_0;
        v11_0 = v10_0;
    } else {
        // two-sweep WENO5/MPWENO5 corner values (1936-2051): family 1 sweeps
        // d1 then d2, family 0 d2 then d1; digit 1 = side of the first
        // sweep, digit 2 = side of the second sweep, in both families.
        using S = Stencil<R, M>;
        R L[6], Q[6], s[6];
#pragma unroll
        for (int mb = 0; mb < 6; ++mb) {
#pragma unroll
            for (int ma = 0; ma < 6; ++ma) s[ma] = cv(ma - 3, mb - 3);
            S::run(s, c, L[mb], Q[mb]);
        }
        S::run(L, c, v00_1, v01_1);
        S::run(Q, c, v10_1, v11_1);
#pragma unroll
        for (int mb = 0; mb < 6; ++mb) {
#pragma unroll
            for (int ma = 0; ma < 6; ++ma) s[ma] = cv(mb - 3, ma - 3);
            S::run(s, c, L[mb], Q[mb]);
        }
        S::run(L, c, v00_0, v01_0);
        S::run(Q, c, v10_0, v11_0);
    }
    // pick per quadrant by the larger one-sided jump (2058-2106)
    R d00 = sgn1(fabs(cv(-2, -1) - cv(-1, -1)) - fabs(cv(-1, -2) - cv(-1, -1)));
    R d10 = sgn1(fabs(cv(1, -1) - cv(0, -1)) - fabs(cv(0, -2) - cv(0, -1)));
    R d01 = sgn1(fabs(cv(-2, 0) - cv(-1, 0)) - fabs(cv(-1, 1) - cv(-1, 0)));
    R d11 = sgn1(fabs(cv(1, 0) - cv(0, 0)) - fabs(cv(0, 1) - cv(0, 0)));
    R v00 = (R(1) + d00) / R(2) * v00_0 + (R(1) - d00) / R(2) * v00_1;
    R v10 = (R(1) + d10) / R(2) * v10_0 + (R(1) - d10) / R(2) * v10_1;
    R v01 = (R(1) + d01) / R(2) * v01_0 + (R(1) - d01) / R(2) * v01_1;
    R v11 = (R(1) + d11) / R(2) * v11_0 + (R(1) - d11) / R(2) * v11_1;
    return R(.25) * (v00 + v01 + v10 + v11);
}

// ------------------------------------------------------------------------
// edge EMF: get_eijk_LFM (mhd3d3.f90:1754-1856).  K = edge direction.
//   K=1 ei(nx0,ny0+1,nz0+1): v1=vy v2=vz, b1=bj (recon z), b2=bk (recon y)
//   K=2 ej(nx0+1,ny0,nz0+1): v1=vz v2=vx, b1=bk (recon x), b2=bi (recon z)
//   K=3 ek(nx0+1,ny0+1,nz0): v1=vx v2=vy, b1=bi (recon y), b2=bj (recon x)
// ------------------------------------------------------------------------
template <typename R, int M, int K, typename TF>
__global__ void __launch_bounds__(kBlock) k_emf(Dims d, Coef<R> c, const double* __restrict__ v1p,
                                               const double* __restrict__ v2p, const double* __restrict__ b1p,
                                               const double* __restrict__ b2p, const double* __restrict__ vf,
                                               TF* __restrict__ e, long nedge) {
    const long t = gtid();
    if (t >= nedge) return;
    const int ii = d.nx0 + (K != 1), jj = d.ny0 + (K != 2);
    int i, j, k;
    unflat(t, ii, jj, i, j, k);
    const int I = d.ng + i, J = d.ng + j, Kc = d.ng + k;
    const int nx = d.nx, ny = d.ny;
    const ptrdiff_t sx = 1, sy = nx, sz = (ptrdiff_t)nx * ny;
    const ptrdiff_t bc = ix(nx, ny, I, J, Kc);
    const ptrdiff_t s1 = K == 1 ? sy : (K == 2 ? sz : sx);
    const ptrdiff_t s2 = K == 1 ? sz : (K == 2 ? sx : sy);
    const R v1av = vav<R, M>(v1p, bc, s1, s2, c);
    const R v2av = vav<R, M>(v2p, bc, s1, s2, c);
    const R sig1 = sgn1(v1av), sig2 = sgn1(v2av);

    ptrdiff_t base1, st1, base2, st2;
    if (K == 1) {  // b1 = bj (nx,ny+1,nz) along z; b2 = bk (nx,ny,nz+1) along y
        base1 = ix(nx, ny + 1, I, J, Kc);
        st1 = (ptrdiff_t)nx * (ny + 1);
        base2 = ix(nx, ny, I, J, Kc);
        st2 = nx;
    } else if (K == 2) {  // b1 = bk along x; b2 = bi (nx+1,ny,nz) along z
        base1 = ix(nx, ny, I, J, Kc);
        st1 = 1;
        base2 = ix(nx + 1, ny, I, J, Kc);
        st2 = (ptrdiff_t)(nx + 1) * ny;
    } else {  // b1 = bi along y; b2 = bj along x
        base1 = ix(nx + 1, ny, I, J, Kc);
        st1 = nx + 1;
        base2 = ix(nx, ny + 1, I, J, Kc);
        st2 = 1;
    }
    R b1l, b1r, b2l, b2r;
    recon_line<R, M>(b1p, base1, st1, c, b1l, b1r);
    recon_line<R, M>(b2p, base2, st2, c, b2l, b2r);
    const R b1u = (R(1) + sig2) / R(2) * b1l + (R(1) - sig2) / R(2) * b1r;
    const R b2u = (R(1) + sig1) / R(2) * b2l + (R(1) - sig1) / R(2) * b2r;
    R dum = -v1av * b2u + v2av * b1u;
    // Alfven resistivity: mean of |v|+|B|/sqrt(rho) over the 4 cells at the edge
    R eta;
    if (K == 1)
        eta = R(.25) * (R(vf[bc - sy - sz]) + R(vf[bc - sz]) + R(vf[bc]) + R(vf[bc - sy]));
    else if (K == 2)
        eta = R(.25) * (R(vf[bc - sx - sz]) + R(vf[bc - sx]) + R(vf[bc]) + R(vf[bc - sz]));
    else
        eta = R(.25) * (R(vf[bc - sx - sy]) + R(vf[bc - sy]) + R(vf[bc]) + R(vf[bc - sx]));
    dum = dum + eta * (b1l + b2r - b1r - b2l);
    e[t] = TF(dum);
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
