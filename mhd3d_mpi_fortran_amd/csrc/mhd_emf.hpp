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

// Every thread marches exactly T edges (the launcher keeps T <= the march
// axis' edge count, and the last run of a line starts early, overlapping its
// predecessor, whose edges it recomputes but does not store).  With no run-length test the unrolled
// steps form one block and the windows and carried values rename from step
// to step instead of being copied at every conditional join.  A fence ends
// each step: the step's result is pinned there (an empty asm that takes it
// in a register), so the compiler can neither sink the step's arithmetic to
// the result's later use nor hoist the next steps' loads above it -- either
// would keep the whole run's loaded footprint live at once (hundreds of
// registers).
template <typename R>
__device__ __forceinline__ void step_fence(R& v) {
    asm volatile("" : "+v"(v));
    __builtin_amdgcn_sched_barrier(0);
}

// corner velocities of T consecutive edges along the march axis
// When the march runs along d2 (K=1, K=3), a step shifts the footprint by
// one column, so seven of PDM's eight column interpolations carry over: they
// ride along in a register window and only the new column is interpolated
// (the same interp8 calls on the same values: bit-identical).
template <typename R, int M, bool MD1, bool UNIT, int T, typename TI>
__device__ __forceinline__ void vav_march(const TI* __restrict__ lv, ptrdiff_t b, ptrdiff_t s1, ptrdiff_t s2,
                                          const Coef<R>& c, R (&out)[T]) {
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
    auto step = [&](int s) {
        R nxt[N];
        if constexpr (PRE) {
            if (s + 1 < T) W.row(lv, ob + osm + (unsigned)(N - 1 + LO) * osm, oso, nxt);
        }
        out[s] = vav_win<R, M>(W, c, g);
        if (s + 1 < T) {
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
        step_fence(out[s]);
    };
    // binary64 PDM (whose 8x8 window is 128 registers) and MPWENO5 run the
    // steps as a loop: unrolled they need more registers than it saves
    // (measured: 0.93 vs 0.66 ms for the binary64 ej march at 256^3)
    if constexpr ((M == PDM && sizeof(R) == 8) || M == MPWENO5) {
#pragma unroll 1
        for (int s = 0; s < T; ++s) step(s);
    } else {
#pragma unroll
        for (int s = 0; s < T; ++s) step(s);
    }
}

// One PDM2 line through the corner (mhd3d3.f90:1142-1166), held as the raw
// differences of its 4-point stencil (fm2, fm1, f0, fp1): rm = fm1 - fm2,
// r0 = f0 - fm1, rp = fp1 - f0, their slopes pdmb*r and the slopes' signs.
// All of PDM2 but the clamp of the interpolated value depends on these
// alone, so a line that the march slides along itself by one cell keeps two
// of its differences with their slopes and signs, and a line across the
// march is handed on whole to the next edge.  The same operations on the
// same values as PDM2, so the results are bit-identical.
template <typename R>
struct PdmLine {
    R rm, r0, rp, dm, d0, dp, sm, s0, sp, lo, hi;
    __device__ __forceinline__ void set(R fm2, R fm1, R f0, R fp1, R pdmb) {
        hi = dmax(fm1, f0);
        lo = dmin(fm1, f0);
        rm = fm1 - fm2;
        r0 = f0 - fm1;
        rp = fp1 - f0;
        dm = pdmb * rm;
        d0 = pdmb * r0;
        dp = pdmb * rp;
        sm = sgn1(dm);
        s0 = sgn1(d0);
        sp = sgn1(dp);
    }
    // the stencil moved one cell along the line; fm1, f0, fp1 are the new one's
    __device__ __forceinline__ void slide(R fm1, R f0, R fp1, R pdmb) {
        hi = dmax(fm1, f0);
        lo = dmin(fm1, f0);
        rm = r0;
        r0 = rp;
        rp = fp1 - f0;
        dm = d0;
        d0 = dp;
        dp = pdmb * rp;
        sm = s0;
        s0 = sp;
        sp = sgn1(dp);
    }
    // PDM2's left and right states about the interpolated value fi
    __device__ __forceinline__ void finish(R fm1, R f0, R fi, R& fl, R& fr) const {
        const R fii = dmax(lo, dmin(fi, hi));
        const R qm = fabs(sm + s0), qp = fabs(s0 + sp);
        const R dfl = fii - fm1, dfr = f0 - fii;
        fl = fii - s0 * dmax(R(0), fabs(dfl) - qm * fabs(dm));
        fr = fii + s0 * dmax(R(0), fabs(dfr) - qp * fabs(dp));
    }
};

// PDM corner velocities of T consecutive edges (get_vav_center2corner,
// mhd3d3.f90:1858-2116) marching along axis a, the other corner axis being b:
// (a, b) = (d2, d1) for K=1 and K=3, (d1, d2) for K=2.
//   * g[m]: the 8-point interpolation along b of the footprint row at march
//     offset m-4; vi = interp8(g).  A step adds one row, so seven of the eight
//     carry over.  Along d1 first, then d2 is interp2_pdm's order; for K=2
//     (a = d1) that order holds only in fp32, which interpolates along d2 first
//     (the same tensor-product value up to rounding); binary64 K=2 keeps the
//     window march (vav_march) for bit parity.
//   * the four PDM2 lines: two across the march (along b, at a = -1 and 0) and
//     two along it (along a, at b = -1 and 0); a step hands the a = 0 line on
//     as the next a = -1 one and slides the other two (PdmLine).
//   * the corner selection (mhd3d3.f90:2064-2099) compares |rm| and |rp| of
//     the lines: |a - b| = |b - a| exactly.
// Only the 4x5 core of the footprint (b = -2..1, a = -1..3) stays in registers;
// the new row's other taps serve only its interpolation.
template <typename R, bool MD1, bool UNIT, int T, typename TI>
__device__ __forceinline__ void vav_march_pdm(const TI* __restrict__ lv, const Buf<TI>& bv, unsigned ob,
                                              unsigned osm, unsigned oso, const Coef<R>& c,
                                              R (&out)[T]) {
    constexpr unsigned E = sizeof(TI);
    // the 8 taps along b (offsets -4..3) of the row at byte offset orow
    auto row = [&](unsigned orow, R* v) {
        if constexpr (UNIT) {
            const TI* p = reinterpret_cast<const TI*>(reinterpret_cast<const char*>(lv) + orow);
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = R(p[q - 4]);
        } else {
            const unsigned o0 = orow - 4u * oso;
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = R(bv.at(o0, (unsigned)q * oso));
        }
    };
    const R pdmb = c.pdmb;
    R g[8], V[5][4];  // V[a+1][b+2]
    PdmLine<R> C0, C1, L0, L1;
    {
        R w[8][8];  // w[a+4][b+4]
#pragma unroll
        for (int a = 0; a < 8; ++a) row(ob + (unsigned)(a - 4) * osm, w[a]);
#pragma unroll
        for (int a = 0; a < 8; ++a) g[a] = interp8(w[a]);
#pragma unroll
        for (int a = 0; a < 5; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) V[a][b] = w[a + 3][b + 2];
        C0.set(w[3][2], w[3][3], w[3][4], w[3][5], pdmb);
        C1.set(w[4][2], w[4][3], w[4][4], w[4][5], pdmb);
        L0.set(w[2][3], w[3][3], w[4][3], w[5][3], pdmb);
        L1.set(w[2][4], w[3][4], w[4][4], w[5][4], pdmb);
    }
    // fp32: the next row is loaded before the current corner velocity is
    // computed, so its latency hides behind that arithmetic
    constexpr bool PRE = sizeof(R) == 4;
#pragma unroll
    for (int s = 0; s < T; ++s) {
        {
            R nxt[8];
            if constexpr (PRE) {
                if (s + 1 < T) row(ob + (unsigned)(s + 4) * osm, nxt);
            }
            const R vi = interp8(g);
            // X(q): PDM2 along d1 at d2 offset q; Y(p): along d2 at d1 offset p
            const PdmLine<R>& Xm = MD1 ? L0 : C0;
            const PdmLine<R>& X0 = MD1 ? L1 : C1;
            const PdmLine<R>& Ym = MD1 ? C0 : L0;
            const PdmLine<R>& Y0 = MD1 ? C1 : L1;
            // corner values cv(o1, o2) at d1 offset o1, d2 offset o2 (-1, 0)
            auto cv = [&](int o1, int o2) { return MD1 ? V[o1 + 1][o2 + 2] : V[o2 + 1][o1 + 2]; };
            R v00_0, v10_0, v00_1, v01_1, v10_1, v11_1, v01_0, v11_0;
            Xm.finish(cv(-1, -1), cv(0, -1), vi, v00_0, v10_0);
            Ym.finish(cv(-1, -1), cv(-1, 0), vi, v00_1, v01_1);
            Y0.finish(cv(0, -1), cv(0, 0), vi, v10_1, v11_1);
            X0.finish(cv(-1, 0), cv(0, 0), vi, v01_0, v11_0);
            const R d00 = sgn1(fabs(Xm.rm) - fabs(Ym.rm));
            const R d10 = sgn1(fabs(Xm.rp) - fabs(Y0.rm));
            const R d01 = sgn1(fabs(X0.rm) - fabs(Ym.rp));
            const R d11 = sgn1(fabs(X0.rp) - fabs(Y0.rp));
            const R v00 = pick(d00, v00_0, v00_1);
            const R v10 = pick(d10, v10_0, v10_1);
            const R v01 = pick(d01, v01_0, v01_1);
            const R v11 = pick(d11, v11_0, v11_1);
            out[s] = R(.25) * (v00 + v01 + v10 + v11);
            if (s + 1 < T) {
                if constexpr (!PRE) row(ob + (unsigned)(s + 4) * osm, nxt);
                // slide the footprint one cell along a
#pragma unroll
                for (int m = 0; m < 7; ++m) g[m] = g[m + 1];
                g[7] = interp8(nxt);
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) V[a][b] = V[a + 1][b];
#pragma unroll
                for (int b = 0; b < 4; ++b) V[4][b] = nxt[b + 2];
                C0 = C1;
                C1.set(V[1][0], V[1][1], V[1][2], V[1][3], pdmb);
                L0.slide(V[0][1], V[1][1], V[2][1], pdmb);
                L1.slide(V[0][2], V[1][2], V[2][2], pdmb);
            }
        }
        step_fence(out[s]);
    }
}

// workgroups of the ei march: 64 x-lanes by 4 consecutive y edges, one run
// of T edges along z
template <int T>
__host__ __device__ __forceinline__ unsigned emf_tiles(int nx0, int ny0, int nz0) {
    return (unsigned)((nx0 + kTileX - 1) / kTileX) * ((ny0 + 1 + kTileN - 1) / kTileN) * ((nz0 + 1 + T - 1) / T);
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

// threads = ii * no * ceil(nm / T) (K=2, 3) or emf_tiles workgroups (K=1);
// lanes run along x
template <typename R, int M, int K, int T, typename TI, typename TF>
__global__ void __launch_bounds__(kBlock) k_emf_march(Dims d, Coef<R> c, EmfIn<TI> in, R ef,
                                                     const double* __restrict__ efa, TF* __restrict__ e,
                                                     long nthreads) {
    const int ii = d.nx0 + (K != 1);
    const int no = K == 3 ? d.nz0 : (K == 1 ? d.ny0 + 1 : d.ny0);
    const int nm = K == 3 ? d.ny0 + 1 : d.nz0 + 1;
    int i, o, mb;
    if constexpr (K == 1) {
        // both corner axes are strided for ei: a workgroup takes 64 x-lanes
        // of 4 consecutive y edges (emf_tiles), whose four wavefronts share
        // 11 of their 32 window rows per step in the CU's L1
        const unsigned blk = xcd_block();
        const unsigned ntx = (ii + kTileX - 1) / kTileX, nto = (no + kTileN - 1) / kTileN;
        const unsigned tx = blk % ntx, r = blk / ntx, to = r % nto;
        mb = (int)(r / nto);
        i = (int)(tx * kTileX + (threadIdx.x & (kTileX - 1))) + 1;
        o = (int)(to * kTileN + threadIdx.x / kTileX) + 1;
        if (i > ii || o > no || (long)mb * T >= nm) return;
    } else {
        const long t = gtid_xcd();
        if (t >= nthreads) return;
        const unsigned r = (unsigned)t / (unsigned)ii;   // index spaces stay below 2^31
        i = (int)((unsigned)t - r * (unsigned)ii) + 1;
        mb = (int)(r / (unsigned)no);
        o = (int)(r - (unsigned)mb * (unsigned)no) + 1;
    }
    // runs of exactly T edges: the last run of a line starts early; it
    // stores only the edges that are its own (fp32 contraction may round an
    // edge differently at another step of the run, and two threads writing
    // one edge would leave whichever came last)
    const int m0 = (mb * T < nm - T ? mb * T : nm - T) + 1;
    const int skip = mb * T + 1 - m0;
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

    // buffer views (array sizes in bytes): cell arrays and the two face arrays
    constexpr unsigned E = sizeof(TI);
    const unsigned ncell = (unsigned)nx * (unsigned)ny * (unsigned)d.nz;
    const unsigned nf1 = K == 1 ? (unsigned)nx * (ny + 1) * d.nz : (K == 2 ? (unsigned)nx * ny * (d.nz + 1)
                                                                           : (unsigned)(nx + 1) * ny * d.nz);
    const unsigned nf2 = K == 1 ? (unsigned)nx * ny * (d.nz + 1) : (K == 2 ? (unsigned)(nx + 1) * ny * d.nz
                                                                           : (unsigned)nx * (ny + 1) * d.nz);
    const Buf<TI> bv1(in.v1, ncell * E), bv2(in.v2, ncell * E), bb1(in.b1, nf1 * E), bb2(in.b2, nf2 * E),
        bvf(in.vf, ncell * E);

    R v1av[T], v2av[T];
    // the window's other axis is x (unit stride) for K=2 and K=3
    if constexpr (M == PDM && !(MD1 && sizeof(R) == 8)) {
        const ptrdiff_t sm = MD1 ? s1 : s2, so = MD1 ? s2 : s1;
        const unsigned ob = (unsigned)bc0 * E, osm = (unsigned)sm * E, oso = (unsigned)so * E;
        vav_march_pdm<R, MD1, K != 1, T>(in.v1, bv1, ob, osm, oso, c, v1av);
        vav_march_pdm<R, MD1, K != 1, T>(in.v2, bv2, ob, osm, oso, c, v2av);
    } else {
        vav_march<R, M, MD1, K != 1, T>(in.v1, bc0, s1, s2, c, v1av);
        vav_march<R, M, MD1, K != 1, T>(in.v2, bc0, s1, s2, c, v2av);
    }

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
    unsigned ob1 = (unsigned)base1 * E, ob2 = (unsigned)base2 * E, obc = (unsigned)bc0 * E;
    const unsigned ost1 = (unsigned)st1 * E, ost2 = (unsigned)st2 * E, omv1 = (unsigned)mv1 * E,
                   omv2 = (unsigned)mv2 * E, omarch = (unsigned)smarch * E;
    const unsigned osx = E, osy = (unsigned)sy * E, osz = (unsigned)sz * E;
    // One of the two face-B reconstructions runs along the march axis (b1 =
    // bj along z for K=1, b2 = bi along z for K=2, b1 = bi along y for K=3):
    // consecutive edges share all but one of its taps, which ride along in
    // registers.  The other one's taps are contiguous (x: b1 for K=2, b2 for
    // K=3) or strided (buffer loads with uniform offsets).  Of the four cells
    // around the edge, two are the previous edge's.
    using S = Stencil<R, M>;
    constexpr bool MA1 = K != 2;
    const Buf<TI>& bma = MA1 ? bb1 : bb2;
    const unsigned ostma = MA1 ? ost1 : ost2;
    R ta[S::n];
    {
        const unsigned o0 = (MA1 ? ob1 : ob2) + (unsigned)S::lo * ostma;
#pragma unroll
        for (int m = 0; m < S::n; ++m) ta[m] = R(bma.at(o0, (unsigned)m * ostma));
    }
    const unsigned ou = K == 1 ? osy : osx, ow = K == 3 ? osy : osz;
    R f_uw = R(bvf.at(obc - ou - ow, 0u)), f_w = R(bvf.at(obc - ou - ow, ou));
    // a step's loads: the taps of the reconstruction across the march, the
    // march one's new tap and the two new cells; fp32 loads them one step
    // ahead, so their latency hides behind the current step's arithmetic
    struct StepTaps {
        R o[S::n];
        R nm, f0, fu;
    };
    auto load_taps = [&](unsigned o1s, unsigned o2s, unsigned ocs, StepTaps& t) {
        if constexpr (K == 1) {
            const unsigned o0 = o2s + (unsigned)S::lo * ost2;
#pragma unroll
            for (int m = 0; m < S::n; ++m) t.o[m] = R(bb2.at(o0, (unsigned)m * ost2));
        } else {
            const TI* q = reinterpret_cast<const TI*>(reinterpret_cast<const char*>(K == 2 ? in.b1 : in.b2) +
                                                      (K == 2 ? o1s : o2s));
#pragma unroll
            for (int m = 0; m < S::n; ++m) t.o[m] = R(q[S::lo + m]);
        }
        t.nm = R(bma.at((MA1 ? o1s : o2s) + (unsigned)(S::lo + S::n - 1) * ostma, 0u));
        t.f0 = R(bvf.at(ocs, 0u));
        t.fu = R(bvf.at(ocs - ou, 0u));
    };
    constexpr bool PREB = sizeof(R) == 4;
    StepTaps cur;
    load_taps(ob1, ob2, obc, cur);
#pragma unroll
    for (int s = 0; s < T; ++s) {
        {
            StepTaps nxt;
            if constexpr (PREB) {
                if (s + 1 < T) load_taps(ob1 + omv1, ob2 + omv2, obc + omarch, nxt);
            }
            if (s > 0) {
#pragma unroll
                for (int m = 0; m < S::n - 1; ++m) ta[m] = ta[m + 1];
                ta[S::n - 1] = cur.nm;
            }
            const R sig1 = sgn1(v1av[s]), sig2 = sgn1(v2av[s]);
            R b1l, b1r, b2l, b2r;
            if constexpr (MA1) {
                S::run(ta, c, b1l, b1r);
                S::run(cur.o, c, b2l, b2r);
            } else {
                S::run(cur.o, c, b1l, b1r);
                S::run(ta, c, b2l, b2r);
            }
            const R b1u = pick(sig2, b1l, b1r);
            const R b2u = pick(sig1, b2l, b2r);
            R dum = -v1av[s] * b2u + v2av[s] * b1u;
            // the four cells around the edge (offsets u and w along its
            // transverse axes, w the march axis): (u-1,w-1) and (u,w-1) are
            // the previous edge's (u-1,w) and (u,w)
            const R f_0 = cur.f0, f_u = cur.fu;
            R eta;
            if (K == 1)
                eta = R(.25) * (f_uw + f_w + f_0 + f_u);
            else if (K == 2)
                eta = R(.25) * (f_uw + f_u + f_0 + f_w);
            else
                eta = R(.25) * (f_uw + f_w + f_0 + f_u);
            f_uw = f_u;
            f_w = f_0;
            dum = dum + eta * (b1l + b2r - b1r - b2l);
            // ef = |C_a||C_b|/det A of an affine metric (E . a_k per unit
            // computational length), exactly 1 on the reference's grid; a
            // general metric gives E . dl per edge (efa)
            TF ev = TF(dum * (efa ? R(efa[eo]) : ef));
            step_fence(ev);
            if (s >= skip) e[eo] = ev;
            if (s + 1 < T) {
                if constexpr (!PREB) load_taps(ob1 + omv1, ob2 + omv2, obc + omarch, nxt);
                cur = nxt;
            }
            ob1 += omv1;
            ob2 += omv2;
            obc += omarch;
            eo += ostr;
        }
    }
}

}  // namespace mhd
