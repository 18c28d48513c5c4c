// Device numerics of the ideal-MHD step: reconstruction stencils, the
// Gaussian-split moments and per-point helpers shared by every kernel.
//
// Each function restates one routine of mhd3d3.f90 (cited file:line) with the
// same operation order.  Built with -ffp-contract=off for the binary64 path,
// the only differences from the CPU restatement are the ROCm erfc/exp
// (ocml, not correctly rounded) against glibc's.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mhd {

enum Recon { PDM = 1, TVD = 2, WENO5 = 3, MPWENO5 = 4 };

// scalar parameters of the scheme (module pars, mhd3d3.f90:87-91)
template <typename R>
struct Coef {
    R gam, gm1, cfl, pdmb, eps, alpha, beta;
    R rspi;  // sqrt(pi) of the binary64 pi = 4*atan(1) (mhd3d3.f90:39), host libm
};

template <typename R>
__device__ __forceinline__ R sq(R x) { return x * x; }

// quotient and square root of the flux formulas: IEEE in binary64 (the
// reference's arithmetic); in fp32 (mixed precision) the hardware
// reciprocal and square root (1 ulp), which skip the range-scaling
// sequence of the correctly rounded forms
template <typename R>
__device__ __forceinline__ R qdiv(R a, R b) { return a / b; }
template <>
__device__ __forceinline__ float qdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R>
__device__ __forceinline__ R qsqrt(R x) { return sqrt(x); }
// division by a constant: binary64 divides; fp32 multiplies by the rounded
// reciprocal (folded at compile time)
template <typename R>
__device__ __forceinline__ R cdiv(R x, double c) { return x / R(c); }
template <>
__device__ __forceinline__ float cdiv<float>(float x, double c) { return x * float(1. / c); }
template <>
__device__ __forceinline__ float qsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// SIGN(1.,x): IEEE sign bit, -0.0 gives -1
template <typename R>
__device__ __forceinline__ R sgn1(R x) { return copysign(R(1), x); }

template <typename R>
__device__ __forceinline__ R dmax(R a, R b) { return a > b ? a : b; }
template <typename R>
__device__ __forceinline__ R dmin(R a, R b) { return a < b ? a : b; }
// fp32: the hardware max/min, one instruction instead of a compare and a
// select; the same value as the comparisons for every pair of numbers but
// +0/-0 (either may come back), and a NaN operand gives the other operand
template <>
__device__ __forceinline__ float dmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
template <>
__device__ __forceinline__ float dmin<float>(float a, float b) { return __builtin_fminf(a, b); }

// (1+d)/2 a + (1-d)/2 b for d = SIGN(1.,x) = +-1 (mhd3d3.f90:1823-1824,
// 2106-2109); in fp32 the select it equals for finite a and b
template <typename R>
__device__ __forceinline__ R pick(R d, R a, R b) { return (R(1) + d) / R(2) * a + (R(1) - d) / R(2) * b; }
template <>
__device__ __forceinline__ float pick<float>(float d, float a, float b) { return d > 0.f ? a : b; }

// interp_pdm 8-point face value, s[0..7] = cells -4..+3 (mhd3d3.f90:871-874)
template <typename R>
__device__ __forceinline__ R interp8(const R* s) {
    return cdiv<R>(R(-3) * s[0] + R(29) * s[1] - R(139) * s[2] + R(533) * s[3] + R(533) * s[4] - R(139) * s[5] +
                       R(29) * s[6] - R(3) * s[7],
                   840.);
}

// PDM2 limiter (mhd3d3.f90:1142-1166)
template <typename R>
__device__ __forceinline__ void pdm2(R fm2, R fm1, R f0, R fp1, R fi, R pdmb, R& fl, R& fr) {
    R maxf = dmax(fm1, f0), minf = dmin(fm1, f0);
    R fii = dmax(minf, dmin(fi, maxf));
    R dfm = pdmb * (fm1 - fm2), df0 = pdmb * (f0 - fm1), dfp = pdmb * (fp1 - f0);
    R sm = sgn1(dfm), s0 = sgn1(df0), sp = sgn1(dfp);
    dfm = fabs(dfm);
    dfp = fabs(dfp);
    R qm = fabs(sm + s0), qp = fabs(s0 + sp);
    R dfl = fii - fm1, dfr = f0 - fii;
    fl = fii - s0 * dmax(R(0), fabs(dfl) - qm * dfm);
    fr = fii + s0 * dmax(R(0), fabs(dfr) - qp * dfp);
}

// TVD van Leer (mhd3d3.f90:1240-1246)
template <typename R>
__device__ __forceinline__ void tvd(R fm2, R fm1, R f0, R fp1, R fi, R eps, R& fl, R& fr) {
    R a = (fm1 - fm2) / (f0 - fm1 + eps);
    R b = (a + fabs(a)) / (R(1) + fabs(a));
    fl = fm1 - b * (fm1 - fi);
    a = (f0 - fp1) / (fm1 - f0 + eps);
    b = (a + fabs(a)) / (R(1) + fabs(a));
    fr = f0 - b * (f0 - fi);
}

// fp32 (mixed precision): the same limiter b(a) = (a+|a|)/(1+|a|) with
// a = n/d written as b = 2n/(n+d) for a > 0 and 0 otherwise.  In fp32 the
// denominator d = f(0) - f(-1) + eps can cancel exactly (eps = 1e-8 is within
// the rounding of velocity differences), where the reference's form gives
// inf/inf; here a = +inf gives its limit b = 2 and 0/0 gives b = 0.
__device__ __forceinline__ float tvd_b(float n, float d) {
    return (n != 0.f && ((n > 0.f) == (d >= 0.f))) ? 2.f * n * __builtin_amdgcn_rcpf(n + d) : 0.f;
}
template <>
__device__ __forceinline__ void tvd<float>(float fm2, float fm1, float f0, float fp1, float fi, float eps, float& fl,
                                           float& fr) {
    fl = fm1 - tvd_b(fm1 - fm2, f0 - fm1 + eps) * (fm1 - fi);
    fr = f0 - tvd_b(f0 - fp1, fm1 - f0 + eps) * (f0 - fi);
}

// one side of WENO5 (mhd3d3.f90:1335-1356); mirrored arguments give the right side
template <typename R>
__device__ __forceinline__ R weno5_side(R fm3, R fm2, R fm1, R f0, R fp1, R eps) {
    R h0 = cdiv<R>(R(2) * fm3 - R(7) * fm2 + R(11) * fm1, 6.);
    R h1 = cdiv<R>(-fm2 + R(5) * fm1 + R(2) * f0, 6.);
    R h2 = cdiv<R>(R(2) * fm1 + R(5) * f0 - fp1, 6.);
    R s0 = R(13. / 12.) * sq(fm3 - R(2) * fm2 + fm1) + R(.25) * sq(fm3 - R(4) * fm2 + R(3) * fm1);
    R s1 = R(13. / 12.) * sq(fm2 - R(2) * fm1 + f0) + R(.25) * sq(fm2 - f0);
    R s2 = R(13. / 12.) * sq(fm1 - R(2) * f0 + fp1) + R(.25) * sq(R(3) * fm1 - R(4) * f0 + fp1);
    s0 = qdiv(R(.1), sq(eps + s0));
    s1 = qdiv(R(.6), sq(eps + s1));
    s2 = qdiv(R(.3), sq(eps + s2));
    R t = s0 + s1 + s2;
    R w0 = qdiv(s0, t), w1 = qdiv(s1, t), w2 = qdiv(s2, t);
    return w0 * h0 + w1 * h1 + w2 * h2;
}

// minmod / median (mhd3d3.f90:1380-1403)
template <typename R>
__device__ __forceinline__ R minmod(R a, R b) {
    return R(.5) * (sgn1(a) + sgn1(b)) * dmin(fabs(a), fabs(b));
}
template <typename R>
__device__ __forceinline__ R median(R a, R b, R c) { return a + minmod(b - a, c - a); }

// MP limiting of one WENO5 side (mhd3d3.f90:1521-1588)
template <typename R>
__device__ __forceinline__ R mp_side(R u, R fm3, R fm2, R fm1, R f0, R fp1, R alpha, R beta) {
    R h0 = fm3 - R(2) * fm2 + fm1;
    R h1 = fm2 - R(2) * fm1 + f0;
    R h2 = fm1 - R(2) * f0 + fp1;
    R s2 = R(4) * h1 - h2;
    R s0 = R(4) * h2 - h1;
    R s1 = minmod(s2, s0);
    s2 = minmod(h1, h2);
    s0 = minmod(s1, s2);
    R w2 = R(4) * h0 - h1;
    R w0 = R(4) * h1 - h0;
    R w1 = minmod(w2, w0);
    w2 = minmod(h0, h1);
    w0 = minmod(w1, w2);
    s1 = fm1 + alpha * (fm1 - fm2);
    s2 = R(.5) * (fm1 + f0 - s0);
    s0 = R(.5) * (R(3) * fm1 - fm2) + cdiv<R>(beta * w0, 3.);
    w1 = dmin(dmin(fm1, f0), s2);
    w2 = dmin(dmin(fm1, s1), s0);
    w0 = dmax(w1, w2);
    h1 = dmax(dmax(fm1, f0), s2);
    h2 = dmax(dmax(fm1, s1), s0);
    h0 = dmin(h1, h2);
    return median(u, w0, h0);
}

// reconstruct_3d on a 1-D stencil s[] (mhd3d3.f90:833-860).  Layout of s:
//   PDM     s[0..7] = cells -4..+3
//   TVD     s[0..3] = cells -2..+1
//   WENO5   s[0..5] = cells -3..+2
template <typename R, int M>
struct Stencil;
template <typename R>
struct Stencil<R, PDM> {
    static constexpr int lo = -4, n = 8;
    __device__ static __forceinline__ void run(const R* s, const Coef<R>& c, R& fl, R& fr) {
        R fi = interp8(s);
        pdm2(s[2], s[3], s[4], s[5], fi, c.pdmb, fl, fr);
    }
};
template <typename R>
struct Stencil<R, TVD> {
    static constexpr int lo = -2, n = 4;
    __device__ static __forceinline__ void run(const R* s, const Coef<R>& c, R& fl, R& fr) {
        R fi = (s[1] + s[2]) / R(2);
        tvd(s[0], s[1], s[2], s[3], fi, c.eps, fl, fr);
    }
};
template <typename R>
struct Stencil<R, WENO5> {
    static constexpr int lo = -3, n = 6;
    __device__ static __forceinline__ void run(const R* s, const Coef<R>& c, R& fl, R& fr) {
        fl = weno5_side(s[0], s[1], s[2], s[3], s[4], c.eps);
        fr = weno5_side(s[5], s[4], s[3], s[2], s[1], c.eps);
    }
};
template <typename R>
struct Stencil<R, MPWENO5> {
    static constexpr int lo = -3, n = 6;
    __device__ static __forceinline__ void run(const R* s, const Coef<R>& c, R& fl, R& fr) {
        R wl = weno5_side(s[0], s[1], s[2], s[3], s[4], c.eps);
        R wr = weno5_side(s[5], s[4], s[3], s[2], s[1], c.eps);
        fl = mp_side(wl, s[0], s[1], s[2], s[3], s[4], c.alpha, c.beta);
        fr = mp_side(wr, s[5], s[4], s[3], s[2], s[1], c.alpha, c.beta);
    }
};

// reconstruct the face between cells base-stride and base of array f
template <typename R, int M, typename T>
__device__ __forceinline__ void recon_line(const T* __restrict__ f, ptrdiff_t base, ptrdiff_t stride,
                                           const Coef<R>& c, R& fl, R& fr) {
    using S = Stencil<R, M>;
    R s[S::n];
#pragma unroll
    for (int m = 0; m < S::n; ++m) s[m] = R(f[base + (ptrdiff_t)(S::lo + m) * stride]);
    S::run(s, c, fl, fr);
}

// 32-bit byte offsets of a stencil's taps around base along stride, shared by
// every array of one element type and layout: loads then take the array's
// uniform base address plus one offset register per tap (mhd_create keeps
// every ghost-inclusive array below 2^32 bytes)
template <typename R, int M, typename T>
struct Taps {
    using S = Stencil<R, M>;
    unsigned o[S::n];
    __device__ __forceinline__ Taps(ptrdiff_t base, ptrdiff_t stride) {
        const unsigned b = (unsigned)base * (unsigned)sizeof(T), s = (unsigned)stride * (unsigned)sizeof(T);
#pragma unroll
        for (int m = 0; m < S::n; ++m) o[m] = b + (unsigned)(S::lo + m) * s;
    }
};

template <typename T>
__device__ __forceinline__ T load_at(const T* __restrict__ f, unsigned byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(f) + byte_off);
}

// raw buffer view of one array below 2^32 bytes: a load takes a per-lane byte
// offset (a VGPR) and a uniform byte offset (an SGPR), so the taps of a
// stencil at a uniform stride from the lane's base cost no address
// arithmetic.  A load past the array's end returns 0 instead of reading
// beyond it.
template <typename T>
struct Buf {
    // descriptor word 3 of gfx9 buffers: 32-bit data format, no swizzle
    static constexpr int kWord3 = 0x00020000;
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const T* p, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), (short)0, (int)bytes, kWord3)) {}
    __device__ __forceinline__ T at(unsigned voff, unsigned soff) const {
        if constexpr (sizeof(T) == 4)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    }
};

// recon_line with precomputed tap offsets
template <typename R, int M, typename T>
__device__ __forceinline__ void recon_taps(const T* __restrict__ f, const Taps<R, M, T>& tp, const Coef<R>& c,
                                           R& fl, R& fr) {
    using S = Stencil<R, M>;
    R s[S::n];
#pragma unroll
    for (int m = 0; m < S::n; ++m) s[m] = R(load_at(f, tp.o[m]));
    S::run(s, c, fl, fr);
}

// recon_line from a byte offset and a byte stride (32-bit); UNIT: the
// stride is one element, the taps are contiguous and load as wide vectors
// from one address
template <typename R, int M, bool UNIT, typename T>
__device__ __forceinline__ void recon_off(const T* __restrict__ f, unsigned ob, unsigned ost, const Coef<R>& c,
                                          R& fl, R& fr) {
    using S = Stencil<R, M>;
    R s[S::n];
    if constexpr (UNIT) {
        const T* p = reinterpret_cast<const T*>(reinterpret_cast<const char*>(f) + ob);
#pragma unroll
        for (int m = 0; m < S::n; ++m) s[m] = R(p[S::lo + m]);
    } else {
        // uniform tap offsets in each load's scalar base: one offset register
        const char* b = reinterpret_cast<const char*>(f);
#pragma unroll
        for (int m = 0; m < S::n; ++m)
            s[m] = R(load_at(reinterpret_cast<const T*>(b + (ptrdiff_t)(S::lo + m) * (ptrdiff_t)ost), ob));
    }
    S::run(s, c, fl, fr);
}

// recon_line through a buffer view: the taps at ob + (lo+m)*ost share one
// offset register, the uniform part going into each load's scalar offset
template <typename R, int M, typename T>
__device__ __forceinline__ void recon_buf(const Buf<T>& f, unsigned ob, unsigned ost, const Coef<R>& c, R& fl,
                                          R& fr) {
    using S = Stencil<R, M>;
    R s[S::n];
    const unsigned o0 = ob + (unsigned)S::lo * ost;  // modulo 2^32: the first tap's offset
#pragma unroll
    for (int m = 0; m < S::n; ++m) s[m] = R(f.at(o0, (unsigned)m * ost));
    S::run(s, c, fl, fr);
}

// get_v_face (mhd3d3.f90:2594-2615): the side's half-moments.
//   side = +1: v0p = .5 erfc(-a v), v1p = v v0p + b
//   side = -1: v0n = .5 erfc( a v), v1n = v v0n - b
template <typename R>
__device__ __forceinline__ void vface(R lam, R v, int side, R rspi, R& v0, R& v1);

template <>
__device__ __forceinline__ void vface<double>(double lam, double v, int side, double rspi, double& v0,
                                              double& v1) {
    double a = sqrt(lam);
    double b = .5 * exp(-lam * (v * v)) / a / rspi;
    if (side > 0) {
        v0 = .5 * erfc(-a * v);
        v1 = v * v0 + b;
    } else {
        v0 = .5 * erfc(a * v);
        v1 = v * v0 - b;
    }
}

// fp32 (mixed precision): erfc by Abramowitz & Stegun 7.1.26,
//   erfc(u) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-u^2), t = 1/(1 + p u), u >= 0,
// |error| <= 1.5e-7 (absolute), below the fp32 flux's own rounding; exp(-u^2)
// with u = a|v| is the exp(-lam v^2) of the thermal term, so one exponential
// serves both.  erfc(-u) = 2 - erfc(u).
template <>
__device__ __forceinline__ void vface<float>(float lam, float v, int side, float rspi, float& v0, float& v1) {
    const float a = __builtin_amdgcn_sqrtf(lam);
    const float x = a * v;
    const float e = __expf(-(x * x));
    const float b = .5f * e * __builtin_amdgcn_rcpf(a * rspi);
    const float u = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.f + .3275911f * u);
    const float ec =
        t * (.254829592f + t * (-.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f)))) * e;
    // side > 0: v0 = .5 erfc(-x); side < 0: v0 = .5 erfc(x)
    const bool pos = side > 0 ? x <= 0.f : x >= 0.f;
    v0 = .5f * (pos ? ec : 2.f - ec);
    v1 = side > 0 ? v * v0 + b : v * v0 - b;
}

}  // namespace mhd
