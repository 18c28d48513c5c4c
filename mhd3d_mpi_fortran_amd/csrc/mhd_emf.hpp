// Edge EMF (get_eijk_LFM, mhd3d3.f90:1754-1856) as a marching kernel.
//
// Lanes run along x (every tap a coalesced load); each thread walks T edges
// along a transverse axis and keeps the corner-velocity footprint of the
// current edge in a register window, so one step loads a single new
// stencil row of the window instead of the whole 2-D footprint.  The
// arithmetic per edge is exactly that of get_vav_center2corner
// (mhd3d3.f90:1858-2116): same interpolations, limiters and selection in
// the same order, so the binary64 result is bit-identical to the
// reference's.
//
//   K=1  ei: edge along x, corner plane (d1,d2)=(y,z), march along z (= d2)
//   K=2  ej: edge along y, corner plane (d1,d2)=(z,x), march along z (= d1)
//   K=3  ek: edge along z, corner plane (d1,d2)=(x,y), march along y (= d2)
#pragma once
#include "mhd_kernels.hpp"

namespace mhd {

// stencil footprint of the corner velocity per method: offsets LO..LO+N-1
// along both corner axes
template <int M>
struct VavFoot;
template <>
struct VavFoot<PDM> {
    static constexpr int N = 8, LO = -4;
};
template <>
struct VavFoot<TVD> {
    static constexpr int N = 4, LO = -2;
};
template <>
struct VavFoot<WENO5> {
    static constexpr int N = 6, LO = -3;
};
template <>
struct VavFoot<MPWENO5> {
    static constexpr int N = 6, LO = -3;
};

// register window w[a][b]: a = offset along the march axis, b = along the
// other corner axis, both LO..LO+N-1 relative to the current edge's base cell
template <typename R, int M, bool MARCH_D1, bool UNIT = false>
struct VavWin {
    static constexpr int N = VavFoot<M>::N, LO = VavFoot<M>::LO;
    R w[N][N];
    // value at offset o1 along d1 and o2 along d2
    __device__ __forceinline__ R cv(int o1, int o2) const { return MARCH_D1 ? w[o1 - LO][o2 - LO] : w[o2 - LO][o1 - LO]; }

    // ob, osm, oso: byte offset of the base cell and byte strides along the
    // march and the other axis (32-bit: every array stays below 2^32 bytes);
    // UNIT: the other axis is x, each row of the window is contiguous and
    // loads as wide vectors from one address
    template <typename TI>
    __device__ __forceinline__ void row(const TI* __restrict__ lv, unsigned orow, unsigned oso, R* out) {
        if constexpr (UNIT) {
            const TI* p = reinterpret_cast<const TI*>(reinterpret_cast<const char*>(lv) + orow);
#pragma unroll
            for (int q = 0; q < N; ++q) out[q] = R(p[q + LO]);
        } else {
            // the column offsets are uniform: they go into each load's
            // scalar base, so a row needs one offset register, not N
            const char* b = reinterpret_cast<const char*>(lv);
#pragma unroll
            for (int q = 0; q < N; ++q)
                out[q] = R(load_at(reinterpret_cast<const TI*>(b + (ptrdiff_t)(q + LO) * (ptrdiff_t)oso), orow));
        }
    }
    template <typename TI>
    __device__ __forceinline__ void load(const TI* __restrict__ lv, unsigned ob, unsigned osm, unsigned oso) {
#pragma unroll
        for (int a = 0; a < N; ++a) row(lv, ob + (unsigned)(a + LO) * osm, oso, w[a]);
    }
    // advance one cell along the march axis (ob is the new base cell)
    template <typename TI>
    __device__ __forceinline__ void shift(const TI* __restrict__ lv, unsigned ob, unsigned osm, unsigned oso) {
#pragma unroll
        for (int a = 0; a < N - 1; ++a)
#pragma unroll
            for (int q = 0; q < N; ++q) w[a][q] = w[a + 1][q];
        row(lv, ob + (unsigned)(N - 1 + LO) * osm, oso, w[N - 1]);
    }
    // advance one cell with a row loaded beforehand
    __device__ __forceinline__ void shift_in(const R* nxt) {
#pragma unroll
        for (int a = 0; a < N - 1; ++a)
#pragma unroll
            for (int q = 0; q < N; ++q) w[a][q] = w[a + 1][q];
#pragma unroll
        for (int q = 0; q < N; ++q) w[N - 1][q] = nxt[q];
    }
};

// PDM: the d1-interpolation of footprint column m2 (offset m2-4 along d2)
template <typename R, bool MD1, bool U>
__device__ __forceinline__ R pdm_col(const VavWin<R, PDM, MD1, U>& W, int m2) {
    R s[8];
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) s[m1] = W.cv(m1 - 4, m2 - 4);
    return interp8(s);
}

// corner velocity of one edge from its window (mhd3d3.f90:1858-2116); for
// PDM, g holds the 8 column interpolations (interp2_pdm's first sweep)
template <typename R, int M, bool MD1, bool U>
__device__ __forceinline__ R vav_win(const VavWin<R, M, MD1, U>& W, const Coef<R>& c, const R* g = nullptr) {
    R v00_0, v10_0, v01_0, v11_0, v00_1, v01_1, v10_1, v11_1;
    if constexpr (M == PDM) {
        const R vi = interp8(g);
        pdm2(W.cv(-2, -1), W.cv(-1, -1), W.cv(0, -1), W.cv(1, -1), vi, c.pdmb, v00_0, v10_0);
        pdm2(W.cv(-1, -2), W.cv(-1, -1), W.cv(-1, 0), W.cv(-1, 1), vi, c.pdmb, v00_1, v01_1);
        pdm2(W.cv(0, -2), W.cv(0, -1), W.cv(0, 0), W.cv(0, 1), vi, c.pdmb, v10_1, v11_1);
        pdm2(W.cv(-2, 0), W.cv(-1, 0), W.cv(0, 0), W.cv(1, 0), vi, c.pdmb, v01_0, v11_0);
    } else if constexpr (M == TVD) {
        const R vi = (W.cv(-1, -1) + W.cv(0, -1) + W.cv(0, 0) + W.cv(-1, 0)) / R(4);
        tvd(W.cv(-2, 0), W.cv(-1, 0), W.cv(0, 0), W.cv(1, 0), vi, c.eps, v00_0, v10_0);
        tvd(W.cv(0, -2), W.cv(0, -1), W.cv(0, 0), W.cv(0, 1), vi, c.eps, v00_1, v01_1);
        v10_1 = v00_1;
        v11_1 = v01_1;
        v01_0 = v00_0;
        v11_0 = v10_0;
    } else {
        using S = Stencil<R, M>;
        R L[6], Q[6], s[6];
#pragma unroll
        for (int mb = 0; mb < 6; ++mb) {
#pragma unroll
            for (int ma = 0; ma < 6; ++ma) s[ma] = W.cv(ma - 3, mb - 3);
            S::run(s, c, L[mb], Q[mb]);
        }
        S::run(L, c, v00_1, v01_1);
        S::run(Q, c, v10_1, v11_1);
#pragma unroll
        for (int mb = 0; mb < 6; ++mb) {
#pragma unroll
            for (int ma = 0; ma < 6; ++ma) s[ma] = W.cv(mb - 3, ma - 3);
            S::run(s, c, L[mb], Q[mb]);
        }
        S::run(L, c, v00_0, v01_0);
        S::run(Q, c, v10_0, v11_0);
    }
    R d00 = sgn1(fabs(W.cv(-2, -1) - W.cv(-1, -1)) - fabs(W.cv(-1, -2) - W.cv(-1, -1)));
    R d10 = sgn1(fabs(W.cv(1, -1) - W.cv(0, -1)) - fabs(W.cv(0, -2) - W.cv(0, -1)));
    R d01 = sgn1(fabs(W.cv(-2, 0) - W.cv(-1, 0)) - fabs(W.cv(-1, 1) - W.cv(-1, 0)));
    R d11 = sgn1(fabs(W.cv(1, 0) - W.cv(0, 0)) - fabs(W.cv(0, 1) - W.cv(0, 0)));
    R v00 = (R(1) + d00) / R(2) * v00_0 + (R(1) - d00) / R(2) * v00_1;
    R v10 = (R(1) + d10) / R(2) * v10_0 + (R(1) - d10) / R(2) * v10_1;
    R v01 = (R(1) + d01) / R(2) * v01_0 + (R(1) - d01) / R(2) * v01_1;
    R v11 = (R(1) + d11) / R(2) * v11_0 + (R(1) - d11) / R(2) * v11_1;
    return R(.25) * (v00 + v01 + v10 + v11);
}

// corner velocities of `cnt` consecutive edges along the march axis
// When the march runs along d2 (K=1, K=3), a step shifts the footprint by
// one column, so seven of PDM's eight column interpolations carry over: they
// ride along in a register window and only the new column is interpolated
// (the same interp8 calls on the same values: bit-identical).
template <typename R, int M, bool MD1, bool UNIT, int T, typename TI>
__device__ __forceinline__ void vav_march(const TI* __restrict__ lv, ptrdiff_t b, ptrdiff_t s1, ptrdiff_t s2,
                                          int cnt, const Coef<R>& c, R (&out)[T]) {
    const ptrdiff_t sm = MD1 ? s1 : s2, so = MD1 ? s2 : s1;
    constexpr unsigned E = sizeof(TI);
    unsigned ob = (unsigned)b * E;
    const unsigned osm = (unsigned)sm * E, oso = (unsigned)so * E;
    VavWin<R, M, MD1, UNIT> W;
    W.load(lv, ob, osm, oso);
    R g[8];
    if constexpr (M == PDM) {
#pragma unroll
        for (int m2 = 0; m2 < 8; ++m2) g[m2] = pdm_col(W, m2);
    }
    // fp32 UNIT windows (the fewest registers): the next step's row is
    // loaded before the current corner velocity is computed, so its latency
    // hides behind that arithmetic (binary64 windows would lose a wave per
    // SIMD to the extra row)
    constexpr int N = VavWin<R, M, MD1, UNIT>::N, LO = VavWin<R, M, MD1, UNIT>::LO;
    constexpr bool PRE = UNIT && sizeof(R) == 4;
#pragma unroll
    for (int s = 0; s < T; ++s) {
        if (s < cnt) {
            R nxt[N];
            if constexpr (PRE) {
                if (s + 1 < cnt) W.row(lv, ob + osm + (unsigned)(N - 1 + LO) * osm, oso, nxt);
            }
            out[s] = vav_win<R, M>(W, c, g);
            if (s + 1 < cnt) {
                ob += osm;
                if constexpr (PRE) W.shift_in(nxt);
                else W.shift(lv, ob, osm, oso);
                if constexpr (M == PDM) {
                    if constexpr (MD1) {
#pragma unroll
                        for (int m2 = 0; m2 < 8; ++m2) g[m2] = pdm_col(W, m2);
                    } else {
#pragma unroll
                        for (int m2 = 0; m2 < 7; ++m2) g[m2] = g[m2 + 1];
                        g[7] = pdm_col(W, 7);
                    }
                }
            }
        }
    }
}

template <typename T>
struct EmfIn {  // one edge direction's inputs
    const T *v1, *v2, *b1, *b2, *vf;
};

template <int K>
struct EmfGeom {
    // K=1 ei (nx0, ny0+1, nz0+1): other axis y, march z
    // K=2 ej (nx0+1, ny0, nz0+1): other axis y, march z
    // K=3 ek (nx0+1, ny0+1, nz0): other axis z, march y
    static constexpr bool MD1 = (K == 2);
};

// threads = ii * no * ceil(nm / T); lanes run along x
template <typename R, int M, int K, int T, typename TI, typename TF>
__global__ void __launch_bounds__(kBlock) k_emf_march(Dims d, Coef<R> c, EmfIn<TI> in, R ef,
                                                     const double* __restrict__ efa, TF* __restrict__ e,
                                                     long nthreads) {
    const long t = gtid_xcd();
    if (t >= nthreads) return;
    const int ii = d.nx0 + (K != 1);
    const int no = K == 3 ? d.nz0 : (K == 1 ? d.ny0 + 1 : d.ny0);
    const int nm = K == 3 ? d.ny0 + 1 : d.nz0 + 1;
    const unsigned r = (unsigned)t / (unsigned)ii;   // index spaces stay below 2^31
    const int i = (int)((unsigned)t - r * (unsigned)ii) + 1;
    const int mb = (int)(r / (unsigned)no);
    const int o = (int)(r - (unsigned)mb * (unsigned)no) + 1;
    const int m0 = mb * T + 1;
    const int cnt = nm - m0 + 1 < T ? nm - m0 + 1 : T;
    // first edge of this thread
    const int j0 = K == 3 ? m0 : o, k0 = K == 3 ? o : m0;
    const int nx = d.nx, ny = d.ny, ng = d.ng;
    const int I = ng + i, J = ng + j0, Kc = ng + k0;
    const ptrdiff_t sx = 1, sy = nx, sz = (ptrdiff_t)nx * ny;
    const ptrdiff_t s1 = K == 1 ? sy : (K == 2 ? sz : sx);
    const ptrdiff_t s2 = K == 1 ? sz : (K == 2 ? sx : sy);
    const ptrdiff_t smarch = K == 3 ? sy : sz;
    const ptrdiff_t bc0 = ix(nx, ny, I, J, Kc);
    constexpr bool MD1 = EmfGeom<K>::MD1;

    R v1av[T], v2av[T];
    // the window's other axis is x (unit stride) for K=2 and K=3
    vav_march<R, M, MD1, K != 1, T>(in.v1, bc0, s1, s2, cnt, c, v1av);
    vav_march<R, M, MD1, K != 1, T>(in.v2, bc0, s1, s2, cnt, c, v2av);

    // face B reconstructions and the Alfven resistivity, edge by edge
    ptrdiff_t base1, st1, base2, st2, mv1, mv2, ostr;
    if (K == 1) {  // b1 = bj (nx,ny+1,nz) along z; b2 = bk (nx,ny,nz+1) along y; e (nx0,ny0+1,nz0+1)
        base1 = ix(nx, ny + 1, I, J, Kc);
        st1 = (ptrdiff_t)nx * (ny + 1);
        mv1 = st1;
        base2 = ix(nx, ny, I, J, Kc);
        st2 = nx;
        mv2 = (ptrdiff_t)nx * ny;
        ostr = (ptrdiff_t)d.nx0 * (d.ny0 + 1);
    } else if (K == 2) {  // b1 = bk along x; b2 = bi (nx+1,ny,nz) along z; e (nx0+1,ny0,nz0+1)
        base1 = ix(nx, ny, I, J, Kc);
        st1 = 1;
        mv1 = (ptrdiff_t)nx * ny;
        base2 = ix(nx + 1, ny, I, J, Kc);
        st2 = (ptrdiff_t)(nx + 1) * ny;
        mv2 = st2;
        ostr = (ptrdiff_t)(d.nx0 + 1) * d.ny0;
    } else {  // b1 = bi along y; b2 = bj along x; e (nx0+1,ny0+1,nz0)
        base1 = ix(nx + 1, ny, I, J, Kc);
        st1 = nx + 1;
        mv1 = st1;
        base2 = ix(nx, ny + 1, I, J, Kc);
        st2 = 1;
        mv2 = nx;
        ostr = d.nx0 + 1;
    }
    ptrdiff_t eo = K == 1 ? ix(d.nx0, d.ny0 + 1, i, j0, k0)
                          : (K == 2 ? ix(d.nx0 + 1, d.ny0, i, j0, k0) : ix(d.nx0 + 1, d.ny0 + 1, i, j0, k0));
    // 32-bit byte offsets (every array stays below 2^32 bytes)
    constexpr unsigned E = sizeof(TI);
    unsigned ob1 = (unsigned)base1 * E, ob2 = (unsigned)base2 * E, obc = (unsigned)bc0 * E;
    const unsigned ost1 = (unsigned)st1 * E, ost2 = (unsigned)st2 * E, omv1 = (unsigned)mv1 * E,
                   omv2 = (unsigned)mv2 * E, omarch = (unsigned)smarch * E;
    const unsigned osx = E, osy = (unsigned)sy * E, osz = (unsigned)sz * E;
#pragma unroll
    for (int s = 0; s < T; ++s) {
        if (s < cnt) {
            const R sig1 = sgn1(v1av[s]), sig2 = sgn1(v2av[s]);
            R b1l, b1r, b2l, b2r;
            // b1 runs along x for K=2, b2 for K=3
            recon_off<R, M, K == 2>(in.b1, ob1, ost1, c, b1l, b1r);
            recon_off<R, M, K == 3>(in.b2, ob2, ost2, c, b2l, b2r);
            const R b1u = (R(1) + sig2) / R(2) * b1l + (R(1) - sig2) / R(2) * b1r;
            const R b2u = (R(1) + sig1) / R(2) * b2l + (R(1) - sig1) / R(2) * b2r;
            R dum = -v1av[s] * b2u + v2av[s] * b1u;
            // the four cells around the edge: uniform offsets in the scalar bases
            auto vfat = [&](unsigned back) {
                return R(load_at(reinterpret_cast<const TI*>(reinterpret_cast<const char*>(in.vf) - (ptrdiff_t)back),
                                 obc));
            };
            R eta;
            if (K == 1)
                eta = R(.25) * (vfat(osy + osz) + vfat(osz) + vfat(0u) + vfat(osy));
            else if (K == 2)
                eta = R(.25) * (vfat(osx + osz) + vfat(osx) + vfat(0u) + vfat(osz));
            else
                eta = R(.25) * (vfat(osx + osy) + vfat(osy) + vfat(0u) + vfat(osx));
            dum = dum + eta * (b1l + b2r - b1r - b2l);
            // ef = |C_a||C_b|/det A of an affine metric (E . a_k per unit
            // computational length), exactly 1 on the reference's grid; a
            // general metric gives E . dl per edge (efa)
            e[eo] = TF(dum * (efa ? R(efa[eo]) : ef));
            ob1 += omv1;
            ob2 += omv2;
            obc += omarch;
            eo += ostr;
        }
    }
}

}  // namespace mhd
