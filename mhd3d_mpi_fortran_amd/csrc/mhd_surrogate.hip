// Learned face reconstruction (no counterpart in the reference): a small MLP
// maps the normalised 5-point stencils of the 8 primitives at a face to the
// face's left/right states; plain PDM (mhd3d3.f90:862-924, 1069-1172) is the
// fallback wherever a smoothness or positivity check fails, and its face
// states are the targets the net is trained on.
//
//   input   x[5q+k] = (f_q(cell j-2+k) - c_q)/s_q, k=0..4, face between j-1, j
//           c_q = (f_q(j-1)+f_q(j))/2, s_q = max(max_k |f_q(j-2+k) - c_q|,
//           1e-3 * physical scale of q: |rho|, |p|, |v|+sqrt(p/rho) or
//           |B|+sqrt(p) from the centres), so variations far below a
//           variable's scale stay small
//   output  y[q] = (L_q - c_q)/s_q, y[8+q] = (R_q - c_q)/s_q; in the solver
//           clamped to [min, max] of f(j-1), f(j), PDM's own bound
//   net     y = lin(x) + gate(x) * MLP(x): MLP 40 -> 32 -> 32 -> 16,
//           activation g(z) = z*sigmoid(1.702 z) (the sigmoid form of GELU),
//           fp32 master weights; per output of variable q, lin = the fixed
//           4th-order interpolation (-x0 + 7 x1 + 7 x2 - x3)/12 of q's window
//           and gate = the largest |second difference| of q's window, so the
//           MLP learns PDM's correction to lin and the correction vanishes
//           on linear data; in the solver, faces whose largest normalised
//           second difference is at most ROUGH_MIN take lin alone
//   parameters, packed: W1[32][40] b1[32] W2[32][32] b2[32] W3[16][32] b3[16]
//
// Kernels: sample collection from the solver's own state (binary64 PDM
// targets), training (per-sample forward/backward, per-weight gradient
// reduction, Adam) and bf16-MFMA inference fused into the fp32 face flux.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mhd_ops.hpp"

using namespace mhd;

namespace sur {

constexpr int NI = 40, NH = 32, NO = 16;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NI, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NO * NH, NPARAM = OFF_B3 + NO;  // 2896
constexpr int NS = NI + NO;  // floats per training sample: inputs, targets
constexpr float GK = 1.702f;

__device__ __forceinline__ float act(float z) { return z / (1.f + expf(-GK * z)); }
__device__ __forceinline__ float act_grad(float z) {
    const float s = 1.f / (1.f + expf(-GK * z));
    return s + GK * z * s * (1.f - s);
}
// inference form: exp2 and the reciprocal straight from the hardware
// (v_exp_f32, v_rcp_f32); the result is rounded to bf16 right after
__device__ __forceinline__ float act_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z * (-GK * 1.4426950408889634f)));
}

// 5-point window of one variable in arithmetic type A: values, centre
// c = (f(j-1)+f(j))/2 and range m = max |w - c|
template <typename A, typename T>
__device__ __forceinline__ void load5(const T* __restrict__ f, ptrdiff_t base, ptrdiff_t st, A* w, A& c, A& m) {
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = A(f[base + (ptrdiff_t)(k - 2) * st]);
    c = A(.5) * (w[1] + w[2]);
    m = A(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) m = fmax(m, fabs(w[k] - c));
}

// centre c = (w1+w2)/2 and range m = max |w - c| of a loaded window
__device__ __forceinline__ void window_stats(const float* w, float& c, float& m) {
    c = .5f * (w[1] + w[2]);
    m = 0.f;
#pragma unroll
    for (int k = 0; k < 5; ++k) m = fmaxf(m, fabsf(w[k] - c));
}

// physical scale of each variable at the face from the window centres
// c = (rho, vx, vy, vz, p, bx, by, bz)
template <typename A>
__device__ __forceinline__ void var_scales(const A* c, A* sc) {
    const A r = fabs(c[0]), p = fabs(c[4]);
    const A v = fmax(fmax(fabs(c[1]), fabs(c[2])), fabs(c[3])) + qsqrt(qdiv(p, fmax(r, A(1e-30))));
    const A b = fmax(fmax(fabs(c[5]), fabs(c[6])), fabs(c[7])) + qsqrt(p);
    sc[0] = r;
    sc[1] = sc[2] = sc[3] = v;
    sc[4] = p;
    sc[5] = sc[6] = sc[7] = b;
}
constexpr float NEGLIGIBLE = 1e-3f;
// below this largest normalised second difference the face takes the linear
// part alone: the gate bounds the net's correction by ROUGH_MIN times the
// window scale there (windows that are constant up to rounding noise, or
// linear to within 1e-3 of their range), so running the MLP buys nothing
constexpr float ROUGH_MIN = 1e-3f;

// normalisation of a window: s = max(range, 1e-3 * physical scale), so that
// variations far below the variable's scale stay small after normalisation
// (they carry no weight in the loss and no roughness)
template <typename A>
__device__ __forceinline__ A norm_scale(A m, A sc) {
    return fmax(fmax(m, A(NEGLIGIBLE) * sc), A(1e-30));
}

// the fixed linear part of the net: 4th-order interpolation of each window
__device__ __forceinline__ float lin4(const float* x) {
    return (-x[0] + 7.f * x[1] + 7.f * x[2] - x[3]) * (1.f / 12.f);
}

// smoothness of a normalised window: largest normalised second difference;
// it is also the gate of the learned correction of that variable's outputs
__device__ __forceinline__ float rough5(const float* x) {
    float r = 0.f;
#pragma unroll
    for (int k = 1; k < 4; ++k) r = fmaxf(r, fabsf(x[k - 1] - 2.f * x[k] + x[k + 1]));
    return r;
}

struct State {
    float* p = nullptr;        // fp32 master parameters
    float* g = nullptr;        // gradient accumulator
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    float* samples = nullptr;  // [cap][NS]
    long cap = 0;
    long* count = nullptr;     // device counter of collected samples
    float* scratch = nullptr;  // per-sample activations/deltas of one minibatch
    long scratch_n = 0;
    float* loss = nullptr;     // device loss accumulator
    unsigned short* wbf = nullptr;  // bf16 MFMA A operands per lane (k_pack_bf16)
    float* bias = nullptr;     // biases in D-register order per half-wave (fp32)
    unsigned long long* stats = nullptr;  // faces evaluated, fallbacks
    int* fix = nullptr;                   // fallback-list segments of the three directions
    int* net = nullptr;                   // net-list segments
    unsigned long long* ctr = nullptr;    // the segments' lengths
    long seg_n = 0;                       // capacity in segments (all directions)
    long* blk = nullptr;                  // sample collection: per-workgroup counts and first slots
    long blk_n = 0;
    long step = 0;
    bool have_weights = false;
};

// the sample of face t along D: the normalised windows x and the PDM face
// values y (targets); false if the face fails the smoothness test.  With
// full = false only x and the test are formed.
template <int D>
__device__ __forceinline__ bool collect_face(const Dims& d, const Coef<double>& c, const Prim<double>& q, float theta,
                                             long t, bool full, float* x, float* y) {
    const int ii = d.nx0 + (D == 1), jj = d.ny0 + (D == 2);
    int i, j, k;
    unflat(t, ii, jj, i, j, k);
    const ptrdiff_t base = ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k);
    const ptrdiff_t st = D == 1 ? 1 : (D == 2 ? (ptrdiff_t)d.nx : (ptrdiff_t)d.nx * d.ny);
    const double* f[8] = {q.rho, q.vx, q.vy, q.vz, q.p, q.bx, q.by, q.bz};
    double wd[NI], cc[8], mm[8], sc[8];
    float rough = 0.f;
#pragma unroll
    for (int v = 0; v < 8; ++v) load5<double>(f[v], base, st, wd + 5 * v, cc[v], mm[v]);
    var_scales<double>(cc, sc);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const double ss = norm_scale<double>(mm[v], sc[v]), r = 1. / ss;
#pragma unroll
        for (int a = 0; a < 5; ++a) x[5 * v + a] = float((wd[5 * v + a] - cc[v]) * r);
        rough = fmaxf(rough, rough5(x + 5 * v));
        if (full) {
            double fl, fr;
            recon_line<double, PDM>(f[v], base, st, c, fl, fr);
            y[v] = float((fl - cc[v]) * r);
            y[8 + v] = float((fr - cc[v]) * r);
        }
    }
    return rough <= theta;
}

// ---------------------------------------------------------------------------
// sample collection: faces of the current (binary64) state along direction D,
// in face order.  Pass 0 counts each workgroup's faces that pass the
// smoothness test, k_collect_scan turns the counts into each workgroup's
// first slot (after the samples already held), pass 1 writes the samples:
// the training set -- and with it the trained net -- is the same from run to
// run (a slot per atomic increment would order the samples, and choose the
// ones kept when the set fills up, by the race).
// ---------------------------------------------------------------------------
template <int D, int PASS>
__global__ void __launch_bounds__(kBlock) k_collect(Dims d, Coef<double> c, Prim<double> q, float theta, int stride,
                                                   int phase, float* __restrict__ out, long cap,
                                                   long* __restrict__ cnt, const long* __restrict__ first,
                                                   long nface) {
    __shared__ int wsum[kBlock / 64];
    const long t = gtid();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    bool keep = t < nface && (t % stride) == phase;
    float x[NI], y[NO];
    if (keep) keep = collect_face<D>(d, c, q, theta, t, PASS == 1, x, y);
    const unsigned long long m = __ballot(keep);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    if constexpr (PASS == 0) {
        if (threadIdx.x == 0) {
            int n = 0;
            for (int q2 = 0; q2 < (int)(blockDim.x >> 6); ++q2) n += wsum[q2];
            cnt[blockIdx.x] = n;
        }
        return;
    } else {
        if (!keep) return;
        long slot = first[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
        for (int q2 = 0; q2 < w; ++q2) slot += wsum[q2];
        if (slot >= cap) return;
        float* o = out + slot * NS;
#pragma unroll
        for (int a = 0; a < NI; ++a) o[a] = x[a];
#pragma unroll
        for (int a = 0; a < NO; ++a) o[NI + a] = y[a];
    }
}

// exclusive prefix of the n workgroup counts after *count (the samples held
// so far) into first[]; *count grows by their total.  One workgroup of 1024.
__global__ void __launch_bounds__(1024) k_collect_scan(const long* __restrict__ cnt, long n, long* __restrict__ first,
                                                      long* __restrict__ count) {
    __shared__ long part[1024];
    const long per = (n + blockDim.x - 1) / blockDim.x, b0 = threadIdx.x * per;
    const long b1 = b0 + per < n ? b0 + per : n;
    long s = 0;
    for (long b = b0; b < b1; ++b) s += cnt[b];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {  // 1024 partial sums, one after the other
        long run = *count;
        for (int q = 0; q < (int)blockDim.x; ++q) {
            const long v = part[q];
            part[q] = run;
            run += v;
        }
        *count = run;
    }
    __syncthreads();
    long run = part[threadIdx.x];
    for (long b = b0; b < b1; ++b) {
        first[b] = run;
        run += cnt[b];
    }
}


// ---------------------------------------------------------------------------
// training
// scratch per sample b: x(40) a1(32) a2(32) dz1(32) dz2(32) dy(16)
// ---------------------------------------------------------------------------
constexpr int SC = NI + 4 * NH + NO;  // 184

// forward + backward of one minibatch; sample b of the batch is sample
// (first + b*mult) mod n of the set (a permutation when gcd(mult, n) = 1)
__global__ void __launch_bounds__(kBlock) k_train_fb(const float* __restrict__ P, const float* __restrict__ set,
                                                    long n, long first, long mult, int batch, float inv_norm,
                                                    float* __restrict__ scratch, float* __restrict__ loss) {
    __shared__ float w[NPARAM];
    for (int a = threadIdx.x; a < NPARAM; a += blockDim.x) w[a] = P[a];
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float l2 = 0.f;
    if (b < batch) {
        const long idx = (long)(((unsigned long long)first + (unsigned long long)b * (unsigned long long)mult) % (unsigned long long)n);
        const float* s = set + idx * NS;
        float* sc = scratch + (long)b * SC;
        float x[NI], z1[NH], a1[NH], z2[NH], a2[NH], y[NO];
        for (int a = 0; a < NI; ++a) {
            x[a] = s[a];
            sc[a] = x[a];
        }
        for (int j = 0; j < NH; ++j) {
            float acc = w[OFF_B1 + j];
            for (int a = 0; a < NI; ++a) acc += w[OFF_W1 + j * NI + a] * x[a];
            z1[j] = acc;
            a1[j] = act(acc);
            sc[NI + j] = a1[j];
        }
        for (int j = 0; j < NH; ++j) {
            float acc = w[OFF_B2 + j];
            for (int a = 0; a < NH; ++a) acc += w[OFF_W2 + j * NH + a] * a1[a];
            z2[j] = acc;
            a2[j] = act(acc);
            sc[NI + NH + j] = a2[j];
        }
        float dy[NO];
        for (int o = 0; o < NO; ++o) {
            float acc = w[OFF_B3 + o];
            for (int a = 0; a < NH; ++a) acc += w[OFF_W3 + o * NH + a] * a2[a];
            const float gt = rough5(x + 5 * (o & 7));
            y[o] = lin4(x + 5 * (o & 7)) + gt * acc;
            const float e = y[o] - s[NI + o];
            l2 += e * e;
            dy[o] = 2.f * e * gt * inv_norm;
            sc[NI + 4 * NH + o] = dy[o];
        }
        float dz2[NH];
        for (int a = 0; a < NH; ++a) {
            float acc = 0.f;
            for (int o = 0; o < NO; ++o) acc += w[OFF_W3 + o * NH + a] * dy[o];
            dz2[a] = acc * act_grad(z2[a]);
            sc[NI + 3 * NH + a] = dz2[a];
        }
        for (int a = 0; a < NH; ++a) {
            float acc = 0.f;
            for (int j = 0; j < NH; ++j) acc += w[OFF_W2 + j * NH + a] * dz2[j];
            sc[NI + 2 * NH + a] = acc * act_grad(z1[a]);
        }
    }
    // block sum of the squared error
    __shared__ float red[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l2;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
        atomicAdd(loss, s);
    }
}

// dParam = sum over the batch of delta_out * activation_in; each workgroup
// covers 64 parameters x 4 batch chunks, grid.y splits the batch further.
// Each chunk's partial sum goes to a slot of its own and k_grad_sum adds the
// kGradChunks slots in chunk order, so the gradient -- and the training --
// does not depend on the order in which the chunks finish.
constexpr int kGradChunks = 64;  // = grid.y (16) x 4 wavefronts
__global__ void __launch_bounds__(kBlock) k_train_grad(const float* __restrict__ scratch, int batch,
                                                      float* __restrict__ gpart) {
    const int pidx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int nch = gridDim.y * 4;
    const int ch = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (pidx >= NPARAM) return;
    int din, ain;  // scratch offsets of delta and input activation (-1: bias)
    if (pidx < OFF_B1) { din = NI + 2 * NH + pidx / NI; ain = pidx % NI; }
    else if (pidx < OFF_W2) { din = NI + 2 * NH + (pidx - OFF_B1); ain = -1; }
    else if (pidx < OFF_B2) { din = NI + 3 * NH + (pidx - OFF_W2) / NH; ain = NI + (pidx - OFF_W2) % NH; }
    else if (pidx < OFF_W3) { din = NI + 3 * NH + (pidx - OFF_B2); ain = -1; }
    else if (pidx < OFF_B3) { din = NI + 4 * NH + (pidx - OFF_W3) / NH; ain = NI + NH + (pidx - OFF_W3) % NH; }
    else { din = NI + 4 * NH + (pidx - OFF_B3); ain = -1; }
    const int per = (batch + nch - 1) / nch;
    const int b0 = ch * per, b1 = b0 + per < batch ? b0 + per : batch;
    float acc = 0.f;
    for (int b = b0; b < b1; ++b) {
        const float* sc = scratch + (long)b * SC;
        acc += sc[din] * (ain >= 0 ? sc[ain] : 1.f);
    }
    gpart[(long)ch * NPARAM + pidx] = acc;
}

__global__ void k_grad_sum(const float* __restrict__ gpart, float* __restrict__ g) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= NPARAM) return;
    float s = 0.f;
    for (int ch = 0; ch < kGradChunks; ++ch) s += gpart[(long)ch * NPARAM + a];
    g[a] = s;
}

__global__ void k_adam(float* __restrict__ P, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1c, float b2c) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= NPARAM) return;
    const float gg = g[a];
    const float mm = .9f * m[a] + .1f * gg;
    const float vv = .999f * v[a] + .001f * gg * gg;
    m[a] = mm;
    v[a] = vv;
    P[a] -= lr * (mm / b1c) / (sqrtf(vv / b2c) + 1e-8f);
    g[a] = 0.f;
}

__global__ void k_zero_f(float* __restrict__ p, int n) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) p[a] = 0.f;
}

// ---------------------------------------------------------------------------
// bf16 MFMA forward, 64 faces per wavefront, transposed form
//
// Each layer computes H^T = W . X^T + b with v_mfma_f32_32x32x16_bf16
// (M=N=32, K=16): A = the weights (32 rows of the layer's outputs), B = the
// layer's inputs for 32 faces (one face per column), so the 64 faces of a
// wave are two column tiles.  Lane layouts of this instruction:
//   A: lane l holds A[l%32][8*(l/32)+v], v=0..7
//   B: lane l holds B[8*(l/32)+v][l%32]
//   D: lane l, register r holds D[drow(r, l/32)][l%32],
//      drow(r, h) = 8*(r/4) + 4*h + r%4
// The D registers r = 8kb..8kb+7 of a lane are exactly the B operand of
// k-block kb of the next layer once the next layer's weight columns are
// permuted to k = 16kb + 8h + v <-> hidden unit drow(8kb + v, h); that
// permutation is applied when the weights are packed, so hidden
// activations never leave the registers.  Inputs and outputs of a face
// live in one lane (lane l owns face l of the wave); the two halves of the
// wave swap the parts the other half's tile needs (one shuffle per packed
// pair).  The packed weights sit in LDS (one copy per block, read as the
// A operands of each layer); activations never leave the registers.
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
// two floats to packed bf16, round to nearest even (v_cvt_pk_bf16_f32)
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned int pack2(float a, float b) {
    bf2 v;
    v.x = (__bf16)a;
    v.y = (__bf16)b;
    return __builtin_bit_cast(unsigned int, v);
}
__host__ __device__ constexpr int drow(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

constexpr int NSLOT = 7;        // A operands per lane: W1 k-blocks 0-2, W2 0-1, W3 0-1
constexpr int NBIAS = 40;       // per half-wave: b1 rows drow(0..15), b2 rows drow(0..15), b3 rows drow(0..7)
constexpr int WBF_N = 64 * NSLOT * 8;

// pack: wbf[(lane*7 + slot)*8 + v] (bf16), bias[h*40 + q] (fp32)
__global__ void k_pack_bf16(const float* __restrict__ P, unsigned short* __restrict__ wbf, float* __restrict__ bias) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < WBF_N) {
        const int v = a & 7, slot = (a >> 3) % NSLOT, l = (a >> 3) / NSLOT;
        const int j = l & 31, h = l >> 5;
        float w = 0.f;
        if (slot < 3) {
            const int k = 16 * slot + 8 * h + v;
            if (k < NI) w = P[OFF_W1 + j * NI + k];
        } else if (slot < 5) {
            w = P[OFF_W2 + j * NH + drow(8 * (slot - 3) + v, h)];
        } else if (j < NO) {
            w = P[OFF_W3 + j * NH + drow(8 * (slot - 5) + v, h)];
        }
        wbf[a] = to_bf16(w);
    }
    if (a < 2 * NBIAS) {
        const int h = a / NBIAS, q = a % NBIAS;
        bias[a] = q < 16 ? P[OFF_B1 + drow(q, h)] : (q < 32 ? P[OFF_B2 + drow(q - 16, h)] : P[OFF_B3 + drow(q - 32, h)]);
    }
}

// the packed net of one block, staged in LDS once and read per layer
struct NetW {
    bf8 a[64][NSLOT];
    float b[2][NBIAS];
};

__device__ __forceinline__ void load_net(NetW& W, const unsigned short* __restrict__ wbf,
                                         const float* __restrict__ bias) {
    const bf8* src = reinterpret_cast<const bf8*>(wbf);
    bf8* dst = &W.a[0][0];
    for (int q = threadIdx.x; q < 64 * NSLOT; q += blockDim.x) dst[q] = src[q];
    for (int q = threadIdx.x; q < 2 * NBIAS; q += blockDim.x) (&W.b[0][0])[q] = bias[q];
    __syncthreads();
}

__device__ __forceinline__ bf8 as_bf8(u4 u) { return __builtin_bit_cast(bf8, u); }

__device__ __forceinline__ unsigned int swap_halves(unsigned int u) { return (unsigned int)__shfl_xor((int)u, 32, 64); }
__device__ __forceinline__ float swap_halves(float f) { return __shfl_xor(f, 32, 64); }

// hidden layer epilogue: activation of D, packed as the next layer's B
__device__ __forceinline__ void hidden_b(const f16v& d, bf8 (&b)[2]) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        u4 u;
#pragma unroll
        for (int p = 0; p < 4; ++p) u[p] = pack2(act_fast(d[8 * kb + 2 * p]), act_fast(d[8 * kb + 2 * p + 1]));
        b[kb] = as_bf8(u);
    }
}

// MLP outputs (without lin and gate) of this lane's face from its 40
// normalised inputs; every lane of the wave must call it
__device__ __forceinline__ void mlp_core(const NetW& W, const float* x, int lane, float* y) {
    const bool hi = (lane >> 5) != 0;
    bf8 t0[3], t1[3];
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
        u4 lo4, hi4;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int k0 = 16 * kb + 2 * p, k1 = k0 + 8;
            lo4[p] = pack2(k0 < NI ? x[k0] : 0.f, k0 + 1 < NI ? x[k0 + 1] : 0.f);
            hi4[p] = pack2(k1 < NI ? x[k1] : 0.f, k1 + 1 < NI ? x[k1 + 1] : 0.f);
        }
        u4 keep, recv;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            keep[p] = hi ? hi4[p] : lo4[p];
            recv[p] = swap_halves(hi ? lo4[p] : hi4[p]);
        }
        t0[kb] = as_bf8(hi ? recv : keep);
        t1[kb] = as_bf8(hi ? keep : recv);
    }
    const bf8* A = W.a[lane];
    const float* B = W.b[lane >> 5];
    f16v d0, d1;
#pragma unroll
    for (int r = 0; r < 16; ++r) d0[r] = d1[r] = B[r];
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
        const bf8 a = A[kb];
        d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, t0[kb], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, t1[kb], d1, 0, 0, 0);
    }
    bf8 h0[2], h1[2];
    hidden_b(d0, h0);
    hidden_b(d1, h1);
#pragma unroll
    for (int r = 0; r < 16; ++r) d0[r] = d1[r] = B[16 + r];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const bf8 a = A[3 + kb];
        d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, h0[kb], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, h1[kb], d1, 0, 0, 0);
    }
    hidden_b(d0, h0);
    hidden_b(d1, h1);
#pragma unroll
    for (int r = 0; r < 16; ++r) d0[r] = d1[r] = r < 8 ? B[32 + r] : 0.f;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const bf8 a = A[5 + kb];
        d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, h0[kb], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, h1[kb], d1, 0, 0, 0);
    }
    // lane l owns face l: tile 0 for the low half, tile 1 for the high half;
    // output rows drow(r, h), r < 8, come from the half h
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const float own = hi ? d1[r] : d0[r];
        const float other = swap_halves(hi ? d0[r] : d1[r]);
        y[drow(r, 0)] = hi ? other : own;
        y[drow(r, 1)] = hi ? own : other;
    }
}

// forward of host-given normalised inputs (tests): n rows of 40 floats
__global__ void __launch_bounds__(kBlock) k_infer(const float* __restrict__ x, long n,
                                                 const unsigned short* __restrict__ wbf,
                                                 const float* __restrict__ bias, float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    __shared__ NetW W;
    load_net(W, wbf, bias);
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float xr[NI], out[NO];
#pragma unroll
    for (int a = 0; a < NI; ++a) xr[a] = row < n ? x[row * NI + a] : 0.f;
    mlp_core(W, xr, lane, out);
    if (row < n)
        for (int o = 0; o < NO; ++o)
            y[row * NO + o] = lin4(xr + 5 * (o & 7)) + rough5(xr + 5 * (o & 7)) * out[o];
}

// ---------------------------------------------------------------------------
// fp32 face flux with the learned reconstruction, in two passes.
//
// Pass 1 (k_flux_sur) loads every face's 8 windows and takes what needs no
// net: the faces whose windows are flat to within ROUGH_MIN (all gates are
// zero there, y = lin) get the clamped linear part and their flux; faces
// that fail the smoothness or positivity check are appended to a PDM
// fallback list (k_flux_fix); faces that use the net are appended to a
// net list.  Pass 2 (k_flux_mlp) runs the bf16 MFMA MLP over the net list,
// 64 listed faces to a wavefront: MFMA needs the whole wavefront, and the
// faces that use the net (about 1 % at 256^3) are scattered thinly over the
// faces, so evaluating them in place would run the net in wavefronts where
// most lanes discard it.  Lists are appended with one atomic per wavefront.
// Atomics on one address serialise, and a tenth of the wavefronts hold a
// net face (a few hundredths a fallback face), so both kinds of list come in
// segments: the faces of kNetSeg consecutive pass-1 workgroups, with room
// for all of them and a counter per list on a 128-byte line of its own.
// The later passes give each segment one workgroup.
// ---------------------------------------------------------------------------
constexpr int SUR_GROUPS = 4;
constexpr int kNetSeg = 16, kNetSegCap = kNetSeg * SUR_GROUPS * kBlock;
// counters of segment s, each on its own 128-byte line: [kCtrStride s + kCtrNet] net list,
// [.. + kCtrFix] fallback list, and (when counting) [.. + kCtrFaces] faces evaluated and
// [.. + kCtrFb] fallbacks, summed into the context's totals at the end of the flux
constexpr int kCtrStride = 64, kCtrNet = 0, kCtrFix = 16, kCtrFaces = 32, kCtrFb = 48;

// the 8 windows (5 taps each, physical units) of face (i, j, k) along D
template <int D>
__device__ __forceinline__ void face_windows(const Dims& d, const float* const* f, int i, int j, int k, float* x) {
    const ptrdiff_t st = D == 1 ? 1 : (D == 2 ? (ptrdiff_t)d.nx : (ptrdiff_t)d.nx * d.ny);
    // 32-bit byte offsets of the 5 taps from each array's uniform base
    // (every array stays below 2^32 bytes, checked at mhd_create)
    const unsigned base = (unsigned)ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k);
    unsigned off[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) off[a] = (base + (unsigned)((a - 2) * st)) * 4u;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        if constexpr (D == 1) {  // contiguous taps: wide loads from one address
#pragma unroll
            for (int a = 0; a < 5; ++a) x[5 * v + a] = f[v][(ptrdiff_t)base + a - 2];
        } else {
#pragma unroll
            for (int a = 0; a < 5; ++a) x[5 * v + a] = load_at(f[v], off[a]);
        }
    }
}

// pass 1's windows in a strided sweep (D = 2, 3), through LDS.  A workgroup
// marches along the sweep over SUR_GROUPS face tiles of one x-column run:
// tile tn holds face planes 4 tn .. 4 tn + 3 and reads sweep rows
// 4 tn .. 4 tn + 7 (row rho = cell ng - 1 + rho: the planes' 5 taps), so
// consecutive tiles share 4 rows.  The rows sit in an LDS ring of 8 (row rho
// in slot rho mod 8), 64 x-columns per variable; the first tile loads all 8
// rows, every later one only its 4 new rows, each row one coalesced load,
// and every face reads its 5 taps per variable from the ring.  Every thread
// of the workgroup calls it, tile after tile (two barriers); columns and rows
// past the block's edges are clamped onto it (their faces are not stored).
template <int D>
__device__ __forceinline__ void ring_windows(const Dims& d, const float* const* f, unsigned tx, unsigned tn,
                                             unsigned o, int ii, unsigned no, bool first, float* x) {
    __shared__ float win[8][8][kTileX];
    const int lane = threadIdx.x & (kTileX - 1), q = threadIdx.x / kTileX;
    const int cx = d.ng + 1 + min((int)(tx * kTileX) + lane, ii - 1);
    const int co = d.ng + 1 + min((int)o, (int)no - 1);
    const int cmax = D == 2 ? d.ny : d.nz;
    __syncthreads();  // the previous tile's reads are done
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h == 0 && !first) continue;  // rows 4 tn .. 4 tn + 3 are in the ring already
        const int row = 4 * (int)tn + q + kTileN * h;
        const int cs = min(d.ng - 1 + row, cmax);
        const unsigned at = (unsigned)(D == 2 ? ix(d.nx, d.ny, cx, cs, co) : ix(d.nx, d.ny, cx, co, cs));
#pragma unroll
        for (int v = 0; v < 8; ++v) win[v][row & 7][lane] = f[v][at];
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int a = 0; a < 5; ++a) x[5 * v + a] = win[v][(4 * (int)tn + q + a) & 7][lane];
}

// per variable, in physical units: the window centre cc and normalisation
// scale ss, the linear face value A = lin of the window (the 4th-order
// interpolation commutes with the affine normalisation: its weights sum to
// 1), the gate times the window scale G = the largest |second difference|
// of the window, and the clamp range [lo, hi] of the two cells next to the
// face (PDM's own bound, mhd3d3.f90:1142-1166).  The smoothness test
// rough <= theta on the normalised window is G <= theta * ss; curved = some
// G > ROUGH_MIN * ss.  Both passes derive these the same way.
__device__ __forceinline__ void face_parts(const float* x, float theta, float* cc, float* ss, float* A, float* G,
                                           float* lo, float* hi, bool& smooth, bool& curved) {
    float mm[8], sc[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) window_stats(x + 5 * v, cc[v], mm[v]);
    var_scales<float>(cc, sc);
    smooth = true;
    curved = false;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        ss[v] = norm_scale<float>(mm[v], sc[v]);
        const float w1 = x[5 * v + 1], w2 = x[5 * v + 2];
        lo[v] = w1 < w2 ? w1 : w2;   // (a NaN window fails the smoothness test)
        hi[v] = w1 < w2 ? w2 : w1;
        G[v] = rough5(x + 5 * v);
        A[v] = lin4(x + 5 * v);
        smooth = smooth && (G[v] <= theta * ss[v]);
        curved = curved || (G[v] > ROUGH_MIN * ss[v]);
    }
}

// append the lanes set in `m` (a ballot) to list[*n ...], one atomic per wave
__device__ __forceinline__ void list_append(unsigned long long m, int lane, long t, int* __restrict__ list,
                                            unsigned long long* __restrict__ n) {
    if (!m) return;
    unsigned long long first = 0;
    const int leader = __ffsll((long long)m) - 1;
    if (lane == leader) first = atomicAdd(n, (unsigned long long)__popcll(m));
    first = __shfl(first, leader, 64);
    if ((m >> lane) & 1ull) list[first + __popcll(m & ((1ull << lane) - 1ull))] = (int)t;
}

// the flux of face t from its (clamped) face values
template <int D, int MET>
__device__ __forceinline__ void sur_face_flux(const Dims& d, const Coef<float>& c, const Normals<float>& nm,
                                              const GMet& gm, int i, int j, int k, const float* lv, const float* rv,
                                              float* __restrict__ F, long t, long nface) {
    if (MET == 2) {
        float n[3];
        gen_normal<float, D>(d, gm, i, j, k, n);
        face_flux_met<float, float>(c, lv, rv, n, F, t, nface);
    } else if (MET == 1) {
        face_flux_met<float, float>(c, lv, rv, nm.n[D - 1], F, t, nface);
    } else {
        face_flux<float, D>(c, lv, rv, F, t, nface);
    }
}

// positivity of density and pressure on both sides
__device__ __forceinline__ bool sur_positive(const float* lv, const float* rv) {
    return lv[0] > 0.f && rv[0] > 0.f && lv[4] > 0.f && rv[4] > 0.f;
}

template <int D, int MET>
__global__ void __launch_bounds__(kBlock) k_flux_sur(Dims d, Coef<float> c, Prim<float> q, Normals<float> nm, GMet gm,
                                                    float theta, float* __restrict__ F, long nface,
                                                    int* __restrict__ fix, int* __restrict__ net,
                                                    unsigned long long* __restrict__ ctr,
                                                    unsigned long long* __restrict__ stats) {
    const int lane = threadIdx.x & 63;
    const float* f[8] = {q.rho, q.vx, q.vy, q.vz, q.p, q.bx, q.by, q.bz};
    const int ii = d.nx0 + (D == 1), jj = d.ny0 + (D == 2), kk = d.nz0 + (D == 3);
    const unsigned lb = xcd_block<kXcdRun / SUR_GROUPS>();
    // strided sweeps: this workgroup's x-column run tx, tile run tnc along the
    // sweep (tiles SUR_GROUPS tnc + grp) and transverse index o
    const unsigned nn = D == 2 ? jj : kk, no = D == 2 ? kk : jj;
    const unsigned ntx = (ii + kTileX - 1) / kTileX, ntn = (nn + kTileN - 1) / kTileN;
    const unsigned ntc = (ntn + SUR_GROUPS - 1) / SUR_GROUPS;
    const unsigned tx = lb % ntx, tnc = lb / ntx % ntc, o = lb / ntx / ntc;
    // this workgroup's list segments and their counters
    int* const sfix = fix + (size_t)(lb / kNetSeg) * kNetSegCap;
    int* const snet = net + (size_t)(lb / kNetSeg) * kNetSegCap;
    unsigned long long* const sc = ctr + (size_t)(lb / kNetSeg) * kCtrStride;
    for (int grp = 0; grp < SUR_GROUPS; ++grp) {
        // groups of 256 faces: flat for the x sweep, face tiles marching along
        // the strided sweeps.  Lanes past the end read the first face's
        // windows (their results are never written), so the loads need no
        // branches.
        const unsigned g = lb * SUR_GROUPS + grp;
        const unsigned tn = tnc * SUR_GROUPS + grp;
        int i, j, k;
        long t;
        bool live;
        if constexpr (D == 1) {
            t = (long)g * blockDim.x + threadIdx.x;
            live = t < nface;
            if (__ballot(live) == 0ull) break;  // wave-uniform: later groups are past the end too
            unflat(live ? t : 0, ii, jj, i, j, k);
        } else {
            // every wavefront stays (ring_windows holds barriers), also one
            // whose plane lies past the sweep's end
            live = tn < ntn && tile_face<D>(tx + ntx * (tn + ntn * o), ii, jj, kk, i, j, k);
            if (!live) i = j = k = 1;
            t = ix(ii, jj, i, j, k);
        }
        float x[NI];
        if constexpr (D == 1) face_windows<D>(d, f, i, j, k, x);
        else ring_windows<D>(d, f, tx, tn, o, ii, no, grp == 0, x);
        float cc[8], ss[8], A[8], G[8], lo[8], hi[8];
        bool smooth, curved;
        face_parts(x, theta, cc, ss, A, G, lo, hi, smooth, curved);
        // faces that use the net (smooth) and are curved beyond ROUGH_MIN go
        // to the net list; the others that use the net take y = 0, the
        // clamped linear part
        const bool need = live && smooth && curved;
        list_append(__ballot(need), lane, t, snet, sc + kCtrNet);
        float lv[8], rv[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) lv[v] = rv[v] = __builtin_amdgcn_fmed3f(lo[v], hi[v], A[v]);
        const bool fb = live && !need && (!smooth || !sur_positive(lv, rv));
        const unsigned long long fm = __ballot(fb);
        list_append(fm, lane, t, sfix, sc + kCtrFix);
        if (stats) {  // counting: per-segment tallies (one address for all would serialise)
            const unsigned long long nl = __ballot(live);
            if (lane == 0) {
                atomicAdd(sc + kCtrFaces, (unsigned long long)__popcll(nl));
                if (fm) atomicAdd(sc + kCtrFb, (unsigned long long)__popcll(fm));
            }
        }
        if (live && !need && !fb) sur_face_flux<D, MET>(d, c, nm, gm, i, j, k, lv, rv, F, t, nface);
    }
}

// pass 2 over one net-list segment of n entries: the workgroup's wavefront
// w takes entries [64 w, 64 w + 64), then strides by the workgroup's 256
// (wave-uniform bounds, so every lane of a wave reaches mlp_core together);
// fallbacks join the same segment's fallback list
template <int D, int MET>
__device__ __forceinline__ void flux_net_dir(const NetW& W, float (*stash)[8][kBlock], const Dims& d,
                                             const Coef<float>& c, const Prim<float>& q, const Normals<float>& nm,
                                             const GMet& gm, const int* __restrict__ net, long n,
                                             int* __restrict__ fix, unsigned long long* __restrict__ nfix,
                                             float* __restrict__ F, long nface,
                                             unsigned long long* __restrict__ nfb) {
    const int lane = threadIdx.x & 63;
    const float* f[8] = {q.rho, q.vx, q.vy, q.vz, q.p, q.bx, q.by, q.bz};
    const int ii = d.nx0 + (D == 1), jj = d.ny0 + (D == 2);
    for (long e0 = 64 * (threadIdx.x >> 6); e0 < n; e0 += kBlock) {
        const long e = e0 + lane;
        const bool live = e < n;
        const long t = net[live ? e : e0];  // lanes past the end repeat the wave's first face
        int i, j, k;
        unflat(t, ii, jj, i, j, k);
        float x[NI];
        face_windows<D>(d, f, i, j, k, x);
        float cc[8], ss[8], A[8], G[8], lo[8], hi[8];
        bool smooth, curved;
        face_parts(x, 0.f, cc, ss, A, G, lo, hi, smooth, curved);
        // A, G and the clamp range wait in LDS (each thread its own column)
        // while the MLP holds the registers
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            stash[0][v][threadIdx.x] = A[v];
            stash[1][v][threadIdx.x] = G[v];
            stash[2][v][threadIdx.x] = lo[v];
            stash[3][v][threadIdx.x] = hi[v];
        }
        // the net's inputs: windows centred on cc and divided by ss
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const float r = __frcp_rn(ss[v]);
#pragma unroll
            for (int a = 0; a < 5; ++a) x[5 * v + a] = (x[5 * v + a] - cc[v]) * r;
        }
        float y[NO];
        mlp_core(W, x, lane, y);
        float lv[8], rv[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const float a = stash[0][v][threadIdx.x], g = stash[1][v][threadIdx.x];
            const float l = stash[2][v][threadIdx.x], h = stash[3][v][threadIdx.x];
            // clamp to [lo, hi] (lo <= hi): the median of the three
            lv[v] = __builtin_amdgcn_fmed3f(l, h, a + g * y[v]);
            rv[v] = __builtin_amdgcn_fmed3f(l, h, a + g * y[8 + v]);
        }
        const bool fb = live && !sur_positive(lv, rv);
        const unsigned long long fm = __ballot(fb);
        list_append(fm, lane, t, fix, nfix);
        if (nfb && fm && lane == 0) atomicAdd(nfb, (unsigned long long)__popcll(fm));
        if (live && !fb) sur_face_flux<D, MET>(d, c, nm, gm, i, j, k, lv, rv, F, t, nface);
    }
}

// the list segments of the three directions (blockIdx.y = direction,
// blockIdx.x = segment in the passes over them)
struct SegLists {
    int* net[3];                 // segments of each direction, kNetSegCap apart
    int* fix[3];
    unsigned long long* ctr[3];  // their counters, kCtrStride apart
    unsigned nseg[3];
    float* F[3];
    long nface[3];
};
template <int MET>
__global__ void __launch_bounds__(kBlock) k_flux_mlp(Dims d, Coef<float> c, Prim<float> q, Normals<float> nm, GMet gm,
                                                    const unsigned short* __restrict__ wbf,
                                                    const float* __restrict__ bias, SegLists sl,
                                                    unsigned long long* __restrict__ stats) {
    __shared__ NetW W;
    __shared__ float stash[4][8][kBlock];
    const int dir = blockIdx.y;
    const unsigned sg = blockIdx.x;
    if (sg >= sl.nseg[dir]) return;  // block-uniform
    unsigned long long* const sc = sl.ctr[dir] + (size_t)sg * kCtrStride;
    const long n = (long)sc[kCtrNet];
    if (n == 0) return;  // block-uniform: an empty segment
    load_net(W, wbf, bias);
    const size_t o = (size_t)sg * kNetSegCap;
    unsigned long long* const nfb = stats ? sc + kCtrFb : nullptr;  // counting fallbacks
    if (dir == 0)
        flux_net_dir<1, MET>(W, stash, d, c, q, nm, gm, sl.net[0] + o, n, sl.fix[0] + o, sc + kCtrFix, sl.F[0],
                             sl.nface[0], nfb);
    else if (dir == 1)
        flux_net_dir<2, MET>(W, stash, d, c, q, nm, gm, sl.net[1] + o, n, sl.fix[1] + o, sc + kCtrFix, sl.F[1],
                             sl.nface[1], nfb);
    else
        flux_net_dir<3, MET>(W, stash, d, c, q, nm, gm, sl.net[2] + o, n, sl.fix[2] + o, sc + kCtrFix, sl.F[2],
                             sl.nface[2], nfb);
}

// PDM flux of the listed fallback faces (fp32, as the plain mixed path)
template <int D, int MET>
__device__ __forceinline__ void flux_fix_dir(const Dims& d, const Coef<float>& c, const Prim<float>& q,
                                             const Normals<float>& nm, const GMet& gm, const int* __restrict__ fix,
                                             long n, float* __restrict__ F, long nface) {
    const int ii = d.nx0 + (D == 1), jj = d.ny0 + (D == 2);
    const ptrdiff_t st = D == 1 ? 1 : (D == 2 ? (ptrdiff_t)d.nx : (ptrdiff_t)d.nx * d.ny);
    for (long e = threadIdx.x; e < n; e += blockDim.x) {
        const long t = fix[e];
        int i, j, k;
        unflat(t, ii, jj, i, j, k);
        const ptrdiff_t base = ix(d.nx, d.ny, d.ng + i, d.ng + j, d.ng + k);
        float l[8], r[8];
        recon_line<float, PDM>(q.rho, base, st, c, l[0], r[0]);
        recon_line<float, PDM>(q.vx, base, st, c, l[1], r[1]);
        recon_line<float, PDM>(q.vy, base, st, c, l[2], r[2]);
        recon_line<float, PDM>(q.vz, base, st, c, l[3], r[3]);
        recon_line<float, PDM>(q.p, base, st, c, l[4], r[4]);
        recon_line<float, PDM>(q.bx, base, st, c, l[5], r[5]);
        recon_line<float, PDM>(q.by, base, st, c, l[6], r[6]);
        recon_line<float, PDM>(q.bz, base, st, c, l[7], r[7]);
        if (MET == 2) {
            float nn[3];
            gen_normal<float, D>(d, gm, i, j, k, nn);
            face_flux_met<float, float>(c, l, r, nn, F, t, nface);
        } else if (MET == 1) {
            face_flux_met<float, float>(c, l, r, nm.n[D - 1], F, t, nface);
        } else {
            face_flux<float, D>(c, l, r, F, t, nface);
        }
    }
}

// the fallback-list segments of the three directions in one launch, one
// workgroup per segment
template <int MET>
__global__ void __launch_bounds__(kBlock) k_flux_fix(Dims d, Coef<float> c, Prim<float> q, Normals<float> nm, GMet gm,
                                                    SegLists sl) {
    const int dir = blockIdx.y;
    const unsigned sg = blockIdx.x;
    if (sg >= sl.nseg[dir]) return;
    const long n = (long)sl.ctr[dir][(size_t)sg * kCtrStride + kCtrFix];
    const size_t o = (size_t)sg * kNetSegCap;
    if (dir == 0) flux_fix_dir<1, MET>(d, c, q, nm, gm, sl.fix[0] + o, n, sl.F[0], sl.nface[0]);
    else if (dir == 1) flux_fix_dir<2, MET>(d, c, q, nm, gm, sl.fix[1] + o, n, sl.F[1], sl.nface[1]);
    else flux_fix_dir<3, MET>(d, c, q, nm, gm, sl.fix[2] + o, n, sl.F[2], sl.nface[2]);
}

// the segments' counters
__global__ void k_zero_ctr(unsigned long long* __restrict__ p, long nseg) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nseg) {
        p[t * kCtrStride + kCtrNet] = 0ull;
        p[t * kCtrStride + kCtrFix] = 0ull;
        p[t * kCtrStride + kCtrFaces] = 0ull;
        p[t * kCtrStride + kCtrFb] = 0ull;
    }
}

// the segments' tallies of one flux evaluation added to the context's totals
// (faces evaluated, fallbacks); one workgroup of 1024
__global__ void __launch_bounds__(1024) k_stats_sum(const unsigned long long* __restrict__ p, long nseg,
                                                   unsigned long long* __restrict__ stats) {
    __shared__ unsigned long long red[2][1024 / 64];
    unsigned long long a = 0, b = 0;
    for (long t = threadIdx.x; t < nseg; t += blockDim.x) {
        a += p[t * kCtrStride + kCtrFaces];
        b += p[t * kCtrStride + kCtrFb];
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        stats[0] += a;
        stats[1] += b;
    }
}


}  // namespace sur

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct SurrogateState : sur::State {};

static SurrogateState* sstate(mhd_ctx* c) {
    if (!c->sur) {
        SurrogateState* s = new SurrogateState();
        const size_t pb = sur::NPARAM * sizeof(float);
        bool ok = hipMalloc(&s->p, pb) == hipSuccess && hipMalloc(&s->g, pb) == hipSuccess &&
                  hipMalloc(&s->m, pb) == hipSuccess && hipMalloc(&s->v, pb) == hipSuccess &&
                  hipMalloc(&s->count, 64) == hipSuccess && hipMalloc(&s->loss, 64) == hipSuccess &&
                  hipMalloc(&s->wbf, sur::WBF_N * sizeof(unsigned short)) == hipSuccess &&
                  hipMalloc(&s->bias, 2 * sur::NBIAS * sizeof(float)) == hipSuccess &&
                  hipMalloc(&s->stats, 64) == hipSuccess;
        if (!ok) {
            delete s;
            return nullptr;
        }
        c->sur = s;
        if (hipMemsetAsync(s->g, 0, pb, c->s) != hipSuccess || hipMemsetAsync(s->m, 0, pb, c->s) != hipSuccess ||
            hipMemsetAsync(s->v, 0, pb, c->s) != hipSuccess || hipMemsetAsync(s->p, 0, pb, c->s) != hipSuccess ||
            hipMemsetAsync(s->count, 0, 64, c->s) != hipSuccess || hipMemsetAsync(s->stats, 0, 64, c->s) != hipSuccess) {
            surrogate_free(c);
            return nullptr;
        }
    }
    return c->sur;
}

void surrogate_free(mhd_ctx* c) {
    SurrogateState* s = c->sur;
    if (!s) return;
    void* ps[] = {s->p, s->g, s->m, s->v, s->samples, s->count, s->scratch, s->loss, s->wbf, s->bias, s->stats,
                  s->fix, s->net, s->ctr, s->blk};
    for (void* p : ps)
        if (p) hipFree(p);
    delete s;
    c->sur = nullptr;
}

// pass-1 workgroups and net-list segments of each direction
struct SurGrid {
    unsigned blocks[3], nseg[3];
    long segs;  // all directions
};
static SurGrid sur_grid(const mhd_ctx* c) {
    const Dims& d = c->d;
    const unsigned sg = sur::SUR_GROUPS;
    SurGrid g;
    g.blocks[0] = (unsigned)((c->nf[0] + sg * kBlock - 1) / (sg * kBlock));
    // strided sweeps: x-column runs x tile runs along the sweep x transverse index
    const unsigned ntx = (d.nx0 + kTileX - 1) / kTileX;
    g.blocks[1] = ntx * ((d.ny0 + 1 + kTileN * sg - 1) / (kTileN * sg)) * d.nz0;
    g.blocks[2] = ntx * ((d.nz0 + 1 + kTileN * sg - 1) / (kTileN * sg)) * d.ny0;
    g.segs = 0;
    for (int a = 0; a < 3; ++a) {
        g.nseg[a] = (g.blocks[a] + sur::kNetSeg - 1) / sur::kNetSeg;
        g.segs += g.nseg[a];
    }
    return g;
}

template <int MET>
static void flux_sur_launch(mhd_ctx* c, SurrogateState* S, Src<float> s) {
    const Coef<float> co = make_coef<float>(c);
    const Normals<float> nm = make_normals<float>(c);
    const Prim<float> q{s.rho, s.vx, s.vy, s.vz, s.p, s.bx, s.by, s.bz};
    const float theta = c->sur_theta;
    unsigned long long* st = c->sur_stats ? S->stats : nullptr;
    const SurGrid sg = sur_grid(c);
    launch(c, MHD_K_SURR, blocks(sg.segs), dim3(kBlock), sur::k_zero_ctr, S->ctr, sg.segs);
    sur::SegLists sl{};
    for (int a = 0, o = 0; a < 3; o += (int)sg.nseg[a], ++a) {
        sl.net[a] = S->net + (size_t)o * sur::kNetSegCap;
        sl.fix[a] = S->fix + (size_t)o * sur::kNetSegCap;
        sl.ctr[a] = S->ctr + (size_t)o * sur::kCtrStride;
        sl.nseg[a] = sg.nseg[a];
        sl.F[a] = (float*)c->flux[a];
        sl.nface[a] = c->nf[a];
    }
    launch(c, MHD_K_FLUX, dim3(sg.blocks[0]), dim3(kBlock), sur::k_flux_sur<1, MET>, c->d, co, q, nm, c->gm, theta,
           sl.F[0], c->nf[0], sl.fix[0], sl.net[0], sl.ctr[0], st);
    launch(c, MHD_K_FLUX, dim3(sg.blocks[1]), dim3(kBlock), sur::k_flux_sur<2, MET>, c->d, co, q, nm, c->gm, theta,
           sl.F[1], c->nf[1], sl.fix[1], sl.net[1], sl.ctr[1], st);
    launch(c, MHD_K_FLUX, dim3(sg.blocks[2]), dim3(kBlock), sur::k_flux_sur<3, MET>, c->d, co, q, nm, c->gm, theta,
           sl.F[2], c->nf[2], sl.fix[2], sl.net[2], sl.ctr[2], st);
    // the net lists, then the fallback lists (the net pass appends to them),
    // one workgroup per segment
    const unsigned mseg = std::max(sg.nseg[0], std::max(sg.nseg[1], sg.nseg[2]));
    launch(c, MHD_K_SURR, dim3(mseg, 3), dim3(kBlock), sur::k_flux_mlp<MET>, c->d, co, q, nm, c->gm,
           (const unsigned short*)S->wbf, (const float*)S->bias, sl, st);
    launch(c, MHD_K_SURR, dim3(mseg, 3), dim3(kBlock), sur::k_flux_fix<MET>, c->d, co, q, nm, c->gm, sl);
    if (st) launch(c, MHD_K_SURR, dim3(1), dim3(1024), sur::k_stats_sum, (const unsigned long long*)S->ctr, sg.segs, st);
}

void launch_flux_surrogate(mhd_ctx* c, Src<float> s) {
    SurrogateState* S = c->sur;
    if (!S || !S->have_weights) {  // no trained net: plain fp32 PDM
        launch_flux_fp32(c, s);
        return;
    }
    const long nsegs = sur_grid(c).segs;
    if (S->seg_n < nsegs) {
        void* old[] = {S->fix, S->net, S->ctr};
        for (void* p : old)
            if (p) hipFree(p);
        S->fix = S->net = nullptr;
        S->ctr = nullptr;
        const size_t lb = (size_t)nsegs * sur::kNetSegCap * sizeof(int);
        if (hipMalloc(&S->fix, lb) != hipSuccess) S->fix = nullptr;
        if (hipMalloc(&S->net, lb) != hipSuccess) S->net = nullptr;
        if (hipMalloc(&S->ctr, (size_t)nsegs * sur::kCtrStride * sizeof(unsigned long long)) != hipSuccess)
            S->ctr = nullptr;
        const bool ok = S->fix && S->net && S->ctr;
        S->seg_n = ok ? nsegs : 0;
        if (!ok) {  // no room for the face lists: plain fp32 PDM
            launch_flux_fp32(c, s);
            return;
        }
    }
    if (c->met.mode == 2) flux_sur_launch<2>(c, S, s);
    else if (c->met.on) flux_sur_launch<1>(c, S, s);
    else flux_sur_launch<0>(c, S, s);
}

static int sync_or(mhd_ctx* c) {
    if (c->launch_error != hipSuccess) {
        c->launch_error = hipSuccess;
        return MHD_ERR_HIP;
    }
    return hipStreamSynchronize(c->s) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

int mhd_surrogate_set_weights(mhd_ctx* c, const float* w, int n) {
    if (!c || !w || n != sur::NPARAM) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    SurrogateState* S = sstate(c);
    if (!S) return MHD_ERR_HIP;
    if (hipMemcpyAsync(S->p, w, n * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess) return MHD_ERR_HIP;
    launch(c, MHD_K_SURR, dim3((sur::WBF_N + 255) / 256), dim3(256), sur::k_pack_bf16, (const float*)S->p,
           S->wbf, S->bias);
    S->have_weights = true;
    return sync_or(c);
}

int mhd_surrogate_get_weights(mhd_ctx* c, float* w, int n) {
    if (!c || !w || n != sur::NPARAM || !c->sur) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    if (hipMemcpyAsync(w, c->sur->p, n * sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    return sync_or(c);
}

int mhd_surrogate_configure(mhd_ctx* c, double theta, int count_fallbacks) {
    // theta < 0: every face takes the PDM fallback
    if (!c || !std::isfinite(theta)) return MHD_ERR_PARAM;
    c->sur_theta = (float)theta;
    c->sur_stats = count_fallbacks != 0;
    return MHD_OK;
}

int mhd_surrogate_stats(mhd_ctx* c, long* faces, long* fallbacks) {
    if (!c) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    unsigned long long h[2] = {0, 0};
    if (c->sur && hipMemcpyAsync(h, c->sur->stats, sizeof(h), hipMemcpyDeviceToHost, c->s) != hipSuccess)
        return MHD_ERR_HIP;
    int rc = sync_or(c);
    if (faces) *faces = (long)h[0];
    if (fallbacks) *fallbacks = (long)h[1];
    if (c->sur) hipMemsetAsync(c->sur->stats, 0, sizeof(h), c->s);
    return rc;
}

int mhd_surrogate_collect(mhd_ctx* c, double theta, int stride, long capacity, long* collected) {
    if (!c || stride < 1 || capacity < 1 || !c->have_geometry) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    SurrogateState* S = sstate(c);
    if (!S) return MHD_ERR_HIP;
    if (S->cap != capacity) {
        if (S->samples) hipFree(S->samples);
        S->samples = nullptr;
        if (hipMalloc(&S->samples, (size_t)capacity * sur::NS * sizeof(float)) != hipSuccess) return MHD_ERR_HIP;
        S->cap = capacity;
        if (hipMemsetAsync(S->count, 0, 64, c->s) != hipSuccess) return MHD_ERR_HIP;
    }
    const Coef<double> co = make_coef<double>(c);
    const Prim<double> q{c->rho, c->vx, c->vy, c->vz, c->pr, c->bx, c->by, c->bz};
    const int phase = (int)(S->step++ % stride);
    // per-workgroup counts and first slots of the largest face family
    const long nblk = std::max(blocks(c->nf[0]).x, std::max(blocks(c->nf[1]).x, blocks(c->nf[2]).x));
    if (S->blk_n < nblk) {
        if (S->blk) hipFree(S->blk);
        S->blk = nullptr;
        S->blk_n = 0;
        if (hipMalloc(&S->blk, 2 * (size_t)nblk * sizeof(long)) != hipSuccess) return MHD_ERR_HIP;
        S->blk_n = nblk;
    }
    long* cnt = S->blk;
    long* first = S->blk + nblk;
    auto dir = [&](auto kc, auto kw, long nf) {
        const dim3 g = blocks(nf);
        launch(c, MHD_K_SURR, g, dim3(kBlock), kc, c->d, co, q, (float)theta, stride, phase, S->samples, S->cap, cnt,
               (const long*)first, nf);
        launch(c, MHD_K_SURR, dim3(1), dim3(1024), sur::k_collect_scan, (const long*)cnt, (long)g.x, first, S->count);
        launch(c, MHD_K_SURR, g, dim3(kBlock), kw, c->d, co, q, (float)theta, stride, phase, S->samples, S->cap, cnt,
               (const long*)first, nf);
    };
    dir(sur::k_collect<1, 0>, sur::k_collect<1, 1>, c->nf[0]);
    dir(sur::k_collect<2, 0>, sur::k_collect<2, 1>, c->nf[1]);
    dir(sur::k_collect<3, 0>, sur::k_collect<3, 1>, c->nf[2]);
    long h = 0;
    if (hipMemcpyAsync(&h, S->count, sizeof(long), hipMemcpyDeviceToHost, c->s) != hipSuccess) return MHD_ERR_HIP;
    int rc = sync_or(c);
    if (collected) *collected = h < S->cap ? h : S->cap;
    return rc;
}

int mhd_surrogate_load_samples(mhd_ctx* c, const float* samples, long n) {
    if (!c || !samples || n < 1) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    SurrogateState* S = sstate(c);
    if (!S) return MHD_ERR_HIP;
    if (S->cap < n) {
        if (S->samples) hipFree(S->samples);
        S->samples = nullptr;
        if (hipMalloc(&S->samples, (size_t)n * sur::NS * sizeof(float)) != hipSuccess) return MHD_ERR_HIP;
        S->cap = n;
    }
    if (hipMemcpyAsync(S->samples, samples, (size_t)n * sur::NS * sizeof(float), hipMemcpyHostToDevice, c->s) !=
        hipSuccess)
        return MHD_ERR_HIP;
    if (hipMemcpyAsync(S->count, &n, sizeof(long), hipMemcpyHostToDevice, c->s) != hipSuccess) return MHD_ERR_HIP;
    return sync_or(c);
}

int mhd_surrogate_fetch_samples(mhd_ctx* c, float* out, long n) {
    if (!c || !out || !c->sur || n < 1 || n > c->sur->cap) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    if (hipMemcpyAsync(out, c->sur->samples, (size_t)n * sur::NS * sizeof(float), hipMemcpyDeviceToHost, c->s) !=
        hipSuccess)
        return MHD_ERR_HIP;
    return sync_or(c);
}

// the minibatch scratch (activations and deltas of batch samples), then the
// gradient's per-chunk partial sums
static int ensure_scratch(mhd_ctx* c, SurrogateState* S, long batch) {
    if (S->scratch_n >= batch) return MHD_OK;
    if (S->scratch) hipFree(S->scratch);
    S->scratch = nullptr;
    S->scratch_n = 0;
    const size_t n = (size_t)batch * sur::SC + (size_t)sur::kGradChunks * sur::NPARAM;
    if (hipMalloc(&S->scratch, n * sizeof(float)) != hipSuccess) return MHD_ERR_HIP;
    S->scratch_n = batch;
    return MHD_OK;
}

static void run_minibatch(mhd_ctx* c, SurrogateState* S, const float* P, const float* set, long n, long first,
                          long mult, int batch) {
    launch(c, MHD_K_SURR, blocks(batch), dim3(kBlock), sur::k_train_fb, P, set, n, first, mult, batch,
           1.f / (float)(batch * sur::NO), S->scratch, S->loss);
    float* gpart = S->scratch + (size_t)S->scratch_n * sur::SC;
    launch(c, MHD_K_SURR, dim3((sur::NPARAM + 63) / 64, sur::kGradChunks / 4), dim3(kBlock), sur::k_train_grad,
           (const float*)S->scratch, batch, gpart);
    launch(c, MHD_K_SURR, dim3((sur::NPARAM + 255) / 256), dim3(256), sur::k_grad_sum, (const float*)gpart, S->g);
}

static long gcd_l(long a, long b) {
    while (b) {
        long t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int mhd_surrogate_train(mhd_ctx* c, int epochs, int batch, double lr, unsigned seed, double* mean_loss) {
    if (!c || epochs < 1 || batch < 64 || !(lr > 0.)) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    SurrogateState* S = sstate(c);
    if (!S) return MHD_ERR_HIP;
    long n = 0;
    if (hipMemcpyAsync(&n, S->count, sizeof(long), hipMemcpyDeviceToHost, c->s) != hipSuccess) return MHD_ERR_HIP;
    if (hipStreamSynchronize(c->s) != hipSuccess) return MHD_ERR_HIP;
    n = n < S->cap ? n : S->cap;
    if (n < batch) return MHD_ERR_PARAM;
    if (ensure_scratch(c, S, batch) != MHD_OK) return MHD_ERR_HIP;
    if (!S->have_weights) {  // deterministic initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in))
        std::vector<float> w(sur::NPARAM, 0.f);
        unsigned long long z = 0x9E3779B97F4A7C15ull * (seed + 1);
        auto rnd = [&]() {
            z += 0x9E3779B97F4A7C15ull;
            unsigned long long r = z;
            r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
            r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
            r ^= r >> 31;
            return (double)(r >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int cnt, int fan) {
            const double a = 1. / std::sqrt((double)fan);
            for (int q = 0; q < cnt; ++q) w[off + q] = (float)((2. * rnd() - 1.) * a);
        };
        fill(sur::OFF_W1, sur::NH * sur::NI, sur::NI);
        fill(sur::OFF_W2, sur::NH * sur::NH, sur::NH);
        fill(sur::OFF_W3, sur::NO * sur::NH, sur::NH);
        // the last layer starts small: the net starts at the linear part
        for (int q = 0; q < sur::NO * sur::NH; ++q) w[sur::OFF_W3 + q] *= .01f;
        if (hipMemcpyAsync(S->p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess)
            return MHD_ERR_HIP;
        S->have_weights = true;
    }
    const long nb = n / batch;
    double tot = 0.;
    unsigned long long z = 0xD1B54A32D192ED03ull ^ seed;
    long adam_t = 0;
    std::vector<float> lsum(1);
    for (int e = 0; e < epochs; ++e) {
        // a fresh permutation per epoch: sample b -> (first + b*mult) mod n
        z = z * 6364136223846793005ull + 1442695040888963407ull;
        long mult = (long)((z >> 17) % (unsigned long long)n) | 1;
        while (gcd_l(mult, n) != 1) mult += 2;
        const long first = (long)((z >> 7) % (unsigned long long)n);
        launch(c, MHD_K_SURR, dim3(1), dim3(64), sur::k_zero_f, S->loss, 1);
        for (long bi = 0; bi < nb; ++bi) {
            run_minibatch(c, S, S->p, S->samples, n, (first + bi * batch * (mult % n)) % n, mult, batch);
            ++adam_t;
            const double b1c = 1. - std::pow(.9, (double)adam_t), b2c = 1. - std::pow(.999, (double)adam_t);
            const double prog = ((double)e + (double)bi / (double)nb) / (double)epochs;
            const double lr_t = lr * .5 * (1. + std::cos(3.141592653589793 * prog));
            launch(c, MHD_K_SURR, dim3((sur::NPARAM + 255) / 256), dim3(256), sur::k_adam, S->p, S->g, S->m, S->v,
                   (float)lr_t, (float)b1c, (float)b2c);
        }
        float h = 0.f;
        if (hipMemcpyAsync(&h, S->loss, sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess) return MHD_ERR_HIP;
        if (hipStreamSynchronize(c->s) != hipSuccess) return MHD_ERR_HIP;
        tot = (double)h / (double)(nb * batch * sur::NO);
    }
    launch(c, MHD_K_SURR, dim3((sur::WBF_N + 255) / 256), dim3(256), sur::k_pack_bf16, (const float*)S->p,
           S->wbf, S->bias);
    if (mean_loss) *mean_loss = tot;
    return sync_or(c);
}

int mhd_surrogate_grad(mhd_ctx* c, const float* w, const float* samples, int n, double* loss, float* grad) {
    if (!c || !w || !samples || !grad || n < 1) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    SurrogateState* S = sstate(c);
    if (!S) return MHD_ERR_HIP;
    if (ensure_scratch(c, S, n) != MHD_OK) return MHD_ERR_HIP;
    // the given weights go to a buffer of their own: the context's net, its
    // Adam moments and its gradient accumulator are left as they are
    float* dset = nullptr;
    if (hipMalloc(&dset, ((size_t)n * sur::NS + 2 * sur::NPARAM) * sizeof(float)) != hipSuccess) return MHD_ERR_HIP;
    float* dw = dset + (size_t)n * sur::NS;
    float* dg = dw + sur::NPARAM;
    float* g_keep = S->g;
    S->g = dg;
    int rc = MHD_OK;
    if (hipMemcpyAsync(dw, w, sur::NPARAM * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess ||
        hipMemcpyAsync(dset, samples, (size_t)n * sur::NS * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess)
        rc = MHD_ERR_HIP;
    if (rc == MHD_OK) {
        launch(c, MHD_K_SURR, dim3(1), dim3(64), sur::k_zero_f, S->loss, 1);
        launch(c, MHD_K_SURR, dim3((sur::NPARAM + 255) / 256), dim3(256), sur::k_zero_f, dg, sur::NPARAM);
        run_minibatch(c, S, dw, dset, n, 0, 1, n);
        float h = 0.f;
        if (hipMemcpyAsync(grad, dg, sur::NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess ||
            hipMemcpyAsync(&h, S->loss, sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
            rc = MHD_ERR_HIP;
        int r2 = sync_or(c);
        if (rc == MHD_OK) rc = r2;
        if (loss) *loss = (double)h / (double)(n * sur::NO);
    }
    S->g = g_keep;
    hipFree(dset);
    return rc;
}

int mhd_surrogate_infer(mhd_ctx* c, const float* x, long n, float* y) {
    if (!c || !x || !y || n < 1 || !c->sur || !c->sur->have_weights) return MHD_ERR_PARAM;
    hipSetDevice(c->p.device);
    SurrogateState* S = c->sur;
    float *dx = nullptr, *dy = nullptr;
    if (hipMalloc(&dx, (size_t)n * sur::NI * sizeof(float)) != hipSuccess) return MHD_ERR_HIP;
    if (hipMalloc(&dy, (size_t)n * sur::NO * sizeof(float)) != hipSuccess) {
        hipFree(dx);
        return MHD_ERR_HIP;
    }
    int rc = MHD_OK;
    if (hipMemcpyAsync(dx, x, (size_t)n * sur::NI * sizeof(float), hipMemcpyHostToDevice, c->s) != hipSuccess)
        rc = MHD_ERR_HIP;
    if (rc == MHD_OK) {
        launch(c, MHD_K_SURR, blocks(n), dim3(kBlock), sur::k_infer, (const float*)dx, n,
               (const unsigned short*)S->wbf, (const float*)S->bias, dy);
        if (hipMemcpyAsync(y, dy, (size_t)n * sur::NO * sizeof(float), hipMemcpyDeviceToHost, c->s) != hipSuccess)
            rc = MHD_ERR_HIP;
        int r2 = sync_or(c);
        if (rc == MHD_OK) rc = r2;
    }
    hipFree(dx);
    hipFree(dy);
    return rc;
}
