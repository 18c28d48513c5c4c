// Launchers of the stencil kernels (face flux, edge EMF), instantiated once
// per arithmetic type in its own translation unit so that each can carry its
// own floating-point contract: binary64 without contraction (reference
// parity), fp32 with FMA contraction (mixed precision).
#pragma once
#include "mhd_emf.hpp"
#include "mhd_launch.hpp"

namespace mhd {

template <typename R, typename TI, typename TF, int M, int MET>
void flux_launch_mm(mhd_ctx* c, Src<TI> s) {
    const Coef<R> co = make_coef<R>(c);
    const Prim<TI> q{s.rho, s.vx, s.vy, s.vz, s.p, s.bx, s.by, s.bz};
    const Normals<R> nm = make_normals<R>(c);
    launch(c, MHD_K_FLUX, blocks(c->nf[0]), dim3(kBlock), k_flux<R, M, 1, TI, TF, MET>, c->d, co, q, nm, c->gm,
           (TF*)c->flux[0], c->nf[0]);
    const Dims& d = c->d;
    // fp32 inputs march kFluxRun tiles per workgroup where that still leaves
    // kFluxRunBlocks workgroups, else one tile each (k_flux)
    const unsigned y8 = face_runs<2>(d.nx0, d.ny0 + 1, d.nz0, kFluxRun);
    const unsigned z8 = face_runs<3>(d.nx0, d.ny0, d.nz0 + 1, kFluxRun);
    const bool fp32 = sizeof(TI) == 4;
    if (fp32 && y8 >= (unsigned)kFluxRunBlocks)
        launch(c, MHD_K_FLUX, dim3(y8), dim3(kBlock), k_flux<R, M, 2, TI, TF, MET, kFluxRun>, c->d, co, q, nm,
               c->gm, (TF*)c->flux[1], c->nf[1]);
    else
        launch(c, MHD_K_FLUX, dim3(face_tiles<2>(d.nx0, d.ny0 + 1, d.nz0)), dim3(kBlock),
               k_flux<R, M, 2, TI, TF, MET>, c->d, co, q, nm, c->gm, (TF*)c->flux[1], c->nf[1]);
    if (fp32 && z8 >= (unsigned)kFluxRunBlocks)
        launch(c, MHD_K_FLUX, dim3(z8), dim3(kBlock), k_flux<R, M, 3, TI, TF, MET, kFluxRun>, c->d, co, q, nm,
               c->gm, (TF*)c->flux[2], c->nf[2]);
    else
        launch(c, MHD_K_FLUX, dim3(face_tiles<3>(d.nx0, d.ny0, d.nz0 + 1)), dim3(kBlock),
               k_flux<R, M, 3, TI, TF, MET>, c->d, co, q, nm, c->gm, (TF*)c->flux[2], c->nf[2]);
}

template <typename R, typename TI, typename TF, int M>
void flux_launch_m(mhd_ctx* c, Src<TI> s) {
    if (c->met.mode == 2) flux_launch_mm<R, TI, TF, M, 2>(c, s);
    else if (c->met.on) flux_launch_mm<R, TI, TF, M, 1>(c, s);
    else flux_launch_mm<R, TI, TF, M, 0>(c, s);
}

template <typename R, typename TI, typename TF>
void flux_launch(mhd_ctx* c, Src<TI> s) {
    switch (c->p.reconop) {
    case PDM: flux_launch_m<R, TI, TF, PDM>(c, s); break;
    case TVD: flux_launch_m<R, TI, TF, TVD>(c, s); break;
    case WENO5: flux_launch_m<R, TI, TF, WENO5>(c, s); break;
    default: flux_launch_m<R, TI, TF, MPWENO5>(c, s); break;
    }
}

// edges handled per thread along the march axis: long runs share the
// sliding stencil windows; a launch whose long-run grid has fewer threads
// than c->emf_fill takes runs of 2 instead, trading window reuse for latency
// hiding.  PDM's ei march (both corner axes strided, its loads the slowest)
// takes runs of 16: at 256^3 0.20 vs 0.22 ms (fp32) and 0.48 vs 0.51 ms
// (binary64), and runs of 32 are slower again (0.23 ms fp32, 0.83 ms
// binary64); for ej and ek runs of 16 measured slower (0.19 vs 0.17 ms
// fp32), and so did runs of 4 (0.17 vs 0.16 ms) and register caps that
// buy a fifth or sixth wavefront per SIMD with spills (EMF class 0.61 and
// 0.72 vs 0.53 ms).
template <int M, int K>
struct EmfSteps {
    static constexpr int T = (M == TVD || (M == PDM && K == 1)) ? 16 : 8;
    static constexpr int T_SHORT = 2;
};

template <typename R, int M, int K, int T, typename TI, typename TF>
void emf_launch_kt(mhd_ctx* c, EmfIn<TI> in) {
    const Dims& d = c->d;
    const long ii = d.nx0 + (K != 1);
    const long no = K == 3 ? d.nz0 : (K == 1 ? d.ny0 + 1 : d.ny0);
    const long nm = K == 3 ? d.ny0 + 1 : d.nz0 + 1;
    const long n = ii * no * ((nm + T - 1) / T);
    const dim3 grid = K == 1 ? dim3(emf_tiles<T>(d.nx0, d.ny0, d.nz0)) : blocks(n);
    launch(c, MHD_K_EMF, grid, dim3(kBlock), k_emf_march<R, M, K, T, TI, TF>, d, make_coef<R>(c), in,
           R(c->met.ef[K - 1]), c->met.mode == 2 ? c->gm.ef[K - 1] : (const double*)nullptr, (TF*)c->emf[K - 1], n);
}

template <typename R, int M, int K, typename TI, typename TF>
void emf_launch_k(mhd_ctx* c, EmfIn<TI> in) {
    constexpr int T = EmfSteps<M, K>::T;
    const Dims& d = c->d;
    const long ii = d.nx0 + (K != 1);
    const long no = K == 3 ? d.nz0 : (K == 1 ? d.ny0 + 1 : d.ny0);
    const long nm = K == 3 ? d.ny0 + 1 : d.nz0 + 1;
    // a run holds exactly T edges (the last one of a line overlaps its
    // predecessor), so long runs need a march axis of at least T edges; it
    // holds at least 2 = T_SHORT
    if (nm >= T && ii * no * ((nm + T - 1) / T) >= c->emf_fill) emf_launch_kt<R, M, K, T, TI, TF>(c, in);
    else emf_launch_kt<R, M, K, EmfSteps<M, K>::T_SHORT, TI, TF>(c, in);
}

// edge k: (v1, v2, b1, b2) = (vy, vz, bj, bk), (vz, vx, bk, bi), (vx, vy, bi, bj);
// with a metric the velocities are the face-normal ones w_d = v . n_d
template <typename R, typename TI, typename TF, int M>
void emf_launch_m(mhd_ctx* c, Src<TI> s) {
    const TI* u[3] = {s.vx, s.vy, s.vz};
    if (c->met.on) {
        u[0] = s.w[0];
        u[1] = s.w[1];
        u[2] = s.w[2];
    }
    emf_launch_k<R, M, 1, TI, TF>(c, EmfIn<TI>{u[1], u[2], s.bj, s.bk, s.vf});
    emf_launch_k<R, M, 2, TI, TF>(c, EmfIn<TI>{u[2], u[0], s.bk, s.bi, s.vf});
    emf_launch_k<R, M, 3, TI, TF>(c, EmfIn<TI>{u[0], u[1], s.bi, s.bj, s.vf});
}

template <typename R, typename TI, typename TF>
void emf_launch(mhd_ctx* c, Src<TI> s) {
    switch (c->p.reconop) {
    case PDM: emf_launch_m<R, TI, TF, PDM>(c, s); break;
    case TVD: emf_launch_m<R, TI, TF, TVD>(c, s); break;
    case WENO5: emf_launch_m<R, TI, TF, WENO5>(c, s); break;
    default: emf_launch_m<R, TI, TF, MPWENO5>(c, s); break;
    }
}

// get_v_face (mhd3d3.f90:2594-2615) elementwise in the arithmetic type R of
// the context's flux (binary64 inputs and outputs)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_vface(long n, const double* __restrict__ lam,
                                                 const double* __restrict__ v, double* __restrict__ v0p,
                                                 double* __restrict__ v0n, double* __restrict__ v1p,
                                                 double* __restrict__ v1n, R rspi) {
    const long t = gtid();
    if (t >= n) return;
    R a0, a1, b0, b1;
    vface<R>(R(lam[t]), R(v[t]), +1, rspi, a0, a1);
    vface<R>(R(lam[t]), R(v[t]), -1, rspi, b0, b1);
    v0p[t] = double(a0);
    v1p[t] = double(a1);
    v0n[t] = double(b0);
    v1n[t] = double(b1);
}

template <typename R>
void vface_launch(mhd_ctx* c, long n, const double* lam, const double* v, double* v0p, double* v0n, double* v1p,
                  double* v1n) {
    launch(c, MHD_K_FLUX, blocks(n), dim3(kBlock), k_vface<R>, n, lam, v, v0p, v0n, v1p, v1n, make_coef<R>(c).rspi);
}

}  // namespace mhd
