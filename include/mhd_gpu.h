/*
 * mhd_gpu.h -- C-ABI of the MI355X ideal-MHD step library (libmhd_gpu.so).
 *
 * Drop-in for the per-step routines of the reference solver mhd3d3.f90
 * (tiegcm/mhd3d_mpi_fortran): a Fortran driver binds these entry points with
 * ISO_C_BINDING (mhd3d_mpi_fortran_amd/fortran/mhd_gpu_mod.f90).  All
 * arguments are plain pointers and sizes; arrays use the reference layout
 * (Fortran column-major, x fastest, ghost-inclusive):
 *
 *   cell arrays  rho,vx,vy,vz,p,bx,by,bz   (nx,ny,nz)       nx = nx0 + 2*ng
 *   face arrays  bi (nx+1,ny,nz)  bj (nx,ny+1,nz)  bk (nx,ny,nz+1)
 *   edge EMFs    ei (nx0,ny0+1,nz0+1) ej (nx0+1,ny0,nz0+1) ek (nx0+1,ny0+1,nz0)
 *   flux/stress  x-faces (nx0+1,ny0,nz0), y-faces (nx0,ny0+1,nz0), z-faces (nx0,ny0,nz0+1)
 *   (module fields, mhd3d3.f90:251-317)
 *
 * One context owns one block (the whole domain) on one GPU and one HIP
 * stream.  Host arrays are borrowed for the duration of a call.  Every entry
 * point returns MHD_OK (0) or a positive error code; the reference stops
 * with a message instead (mhd3d3.f90:3139-3155, 3596).  Entry points that
 * touch host arrays are synchronous; the others are ordered on the
 * context's stream and complete before the next host-visible result.
 */
#ifndef MHD_GPU_H
#define MHD_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhd_ctx mhd_ctx;

enum {
    MHD_OK = 0,
    MHD_ERR_HIP = 1,       /* a HIP runtime call failed */
    MHD_ERR_PARAM = 2,     /* invalid parameter or size */
    MHD_ERR_STATE = 3,     /* rho<0, NaN, or p<=0 found by mhd_check_state */
    MHD_ERR_UNSUPPORTED = 4
};

enum { MHD_BC_EVEN = 0, MHD_BC_PERIODIC = 2 };
enum { MHD_FP64 = 0, MHD_MIXED = 1 };

/* scheme parameters: module pars (mhd3d3.f90:33-248) and the namelist deck
 * groups &grid &comset &free (mhd3d3.f90:152-156) */
typedef struct {
    int nx0, ny0, nz0;   /* interior cells of the block (whole domain on one GPU) */
    int ng;              /* ghost depth = nord/2; PDM needs 4, WENO5/MPWENO5 3, TVD 2 */
    double gam, cfl, pdmb, eps, alpha, beta;
    int reconop;         /* 1 PDM, 2 TVD (van Leer), 3 WENO5, 4 MPWENO5 */
    int stepop;          /* 1 Adams-Bashforth, 2 TVD RK2, 3 TVD RK3 */
    int bc[3];           /* per axis: MHD_BC_EVEN (reference) or MHD_BC_PERIODIC */
    int precision;       /* MHD_FP64, or MHD_MIXED: fp32 stencils/fluxes/EMF,
                            fp64 conserved accumulation and CT update */
    int surrogate;       /* 0, or 1: learned face reconstruction for the flux
                            (needs mhd_surrogate_set_weights), PDM fallback */
    int device;          /* HIP device ordinal */
} mhd_params;

/* fill p with the reference deck defaults (mhd3d3.in: cfl=.3 pdmb=4 eps=1e-8
 * alpha=2 beta=4, gam=5/3, nord=8, PDM, RK2, even reflection) */
void mhd_default_params(mhd_params *p);

/* create/destroy a context; allocates all device state (reference:
 * alloc_pars + alloc_fields, mhd3d3.f90:222-237, 285-317) */
int mhd_create(mhd_ctx **ctx, const mhd_params *p);
void mhd_destroy(mhd_ctx *ctx);
const char *mhd_error_string(int code);

/* metric dxblock/dyblock/dzblock per index, lengths nx, ny, nz
 * (init_geom, mhd3d3.f90:3483-3485; tensor-product grid) */
int mhd_set_geometry(mhd_ctx *ctx, const double *dx, const double *dy, const double *dz);

/* affine metric (no counterpart in the reference, whose grid is the tensor
 * product of init_geom, mhd3d3.f90:3483-3485): physical x = A x_comp with
 * A (row-major 3x3, det A > 0) applied to the tensor-product grid of
 * mhd_set_geometry -- e.g. a sheared box.  Fluxes are taken along the face
 * normals, constrained transport and the CFL use the faces' areas and the
 * cells' volume; vx,vy,vz and the cell B are physical Cartesian components,
 * bi,bj,bk the normal components B.n of the three face families.  Calling it
 * switches the context to the metric kernels (also for A = I). */
int mhd_set_metric(mhd_ctx *ctx, const double *A);

/* general curvilinear metric (no counterpart in the reference, whose grid is
 * the tensor product of init_geom; for grids such as a spherical shell
 * given by their node coordinates): X, Y, Z hold the physical coordinates of the
 * (nx+1)(ny+1)(nz+1) nodes, column-major, node (i,j,k) the lower corner of
 * cell (i,j,k), ghost cells included.  Per-face unit normals and areas,
 * per-cell volumes, cell-B coefficients and CFL spacings, and per-edge EMF
 * factors are computed once on the host and kept on the device.  Same
 * variable meaning as mhd_set_metric; an affine grid given by its nodes
 * reproduces mhd_set_metric's run.  MHD_ERR_PARAM for a degenerate or
 * inverted cell, face or edge.  Requires mhd_set_geometry first (the index
 * space); the spacings given there are not used afterwards. */
int mhd_set_metric_nodes(mhd_ctx *ctx, const double *X, const double *Y, const double *Z);

/* the general metric's arrays as mhd_set_metric_nodes builds them, on the
 * host (no device call; no counterpart in the reference).  out == NULL:
 * returns the number of doubles needed.  Otherwise fills out, in order: for
 * each face family f = i, j, k the unit normals (3 components, each over the
 * ghost-inclusive face layout), the areas and their reciprocals; for each
 * edge direction the EMF factor per owned edge; then per cell (ghost-
 * inclusive) 1/V, the 18 cell-B coefficients (face lo/hi of i, j, k times
 * x, y, z), the 9 cell normals and the 3 CFL factors, each component a
 * contiguous array.  Returns that count, or -MHD_ERR_PARAM for a bad size
 * or a degenerate or inverted cell, face or edge. */
long mhd_metric_nodes_host(int nx0, int ny0, int nz0, int ng, const double *X, const double *Y, const double *Z,
                           double *out);

/* copy the primitive state in / out (full ghost-inclusive arrays).  Any of
 * the download pointers may be NULL. */
int mhd_upload_state(mhd_ctx *ctx, const double *rho, const double *vx, const double *vy, const double *vz,
                     const double *p, const double *bi, const double *bj, const double *bk);
int mhd_download_state(mhd_ctx *ctx, double *rho, double *vx, double *vy, double *vz, double *p, double *bi,
                       double *bj, double *bk, double *bx, double *by, double *bz);

/* initialization after init_ic (mhd3d3.f90:3327-3377): ghost fill, get_bxyz
 * and, for Adams-Bashforth, get_dt then tau=3e-4, tau0=tau, xp=x.
 * *tau: in = deck tau (read_input sets tau0=tau), out = tau after set-up. */
int mhd_initialize(mhd_ctx *ctx, double *tau);

/* nsteps calls of `step` (mhd3d3.f90:3585-3600): step_ab | step_rk2 | step_rk3.
 * RK: *tau in = deck tau (used every step, as the reference does).
 * AB: tau comes from the CFL reduction each step.  On return *tau is the
 * tau of the last step and *tau_cfl the CFL dt computed at its start (the
 * reference warns when tau_cfl < tau0 for RK, mhd3d3.f90:3658-3661). */
int mhd_step(mhd_ctx *ctx, int nsteps, double *tau, double *tau_cfl);

/* batched ensemble of independent blocks (no counterpart in the reference,
 * which runs one domain per MPI job): n contexts of any sizes and
 * precisions on one device, stepped together by one HIP graph launch per
 * step whose per-member branches run concurrently.  Members must use a
 * Runge-Kutta step (stepop 2 or 3: in place, fixed tau), no learned
 * reconstruction, and be initialised (mhd_initialize); tau[m] is member m's
 * fixed step, as *tau of mhd_step.  The graph is recorded once, from the
 * members' launch sequences at creation.  Members stay usable on their
 * own between ensemble steps (upload, download, mhd_step); a member must
 * not be destroyed before the ensemble.  MHD_ERR_PARAM / MHD_ERR_UNSUPPORTED
 * for an empty or mixed-device set or an unsupported member. */
typedef struct mhd_ens mhd_ens;
int mhd_ensemble_create(mhd_ens **ens, mhd_ctx *const *members, int n, const double *tau);
int mhd_ensemble_step(mhd_ens *ens, int nsteps);
void mhd_ensemble_destroy(mhd_ens *ens);

/* per-routine drop-ins.  With them a caller runs the reference's step
 * routine by routine, e.g. step_rk2 (mhd3d3.f90:3648-3711):
 *     mhd_get_dt; mhd_set_tau(tau0, tau0);
 *     for stage 1, 2: mhd_eijk_lfm(CURRENT); mhd_flux_stress(CURRENT);
 *                     mhd_advance(stage); mhd_fill_ghosts; mhd_get_bxyz
 * and step_ab (3602-3644):
 *     mhd_get_dt; mhd_predict_ab; mhd_eijk_lfm(PREDICTED);
 *     mhd_flux_stress(PREDICTED); mhd_advance(0); mhd_fill_ghosts; mhd_get_bxyz
 * which reproduces mhd_step bit for bit. */

/* get_dt, mhd3d3.f90:1608-1639: *tau_cfl = the CFL step, min over the
 * ghost-inclusive block; like the reference it also sets the context's tau */
int mhd_get_dt(mhd_ctx *ctx, double *tau_cfl);
int mhd_fill_ghosts(mhd_ctx *ctx);                /* fill_hydro_mpi + fill_mag_mpi, 781-819 */
int mhd_get_bxyz(mhd_ctx *ctx);                   /* get_bxyz, 2118 */

/* the state a routine reads: the reference passes it as the array
 * arguments of get_flux_stress / get_eijk_LFM -- the current state in
 * step_rk2/rk3 (mhd3d3.f90:3665-3667), the Adams-Bashforth extrapolation
 * rhoh, vxh, ... in step_ab (3618-3620).  PREDICTED needs an
 * Adams-Bashforth context whose mhd_predict_ab is the last routine that
 * changed the state (MHD_ERR_PARAM otherwise). */
enum { MHD_STATE_CURRENT = 0, MHD_STATE_PREDICTED = 1 };
int mhd_flux_stress(mhd_ctx *ctx, int which_state);   /* get_flux_stress, 2385-2560 */
int mhd_eijk_lfm(mhd_ctx *ctx, int which_state);      /* get_eijk_LFM, 1754-1856 */

/* the reference's dead alternatives (their calls, mhd3d3.f90:3354 and 3621,
 * are commented out; the second has none):
 * get_eijk_Yee (1654-1752, with get_v_center 2562-2592): the EMF from the
 * cell-centred electric field split by the Gaussian half-moments and
 * reconstructed face then edge (reconstruct_3d directions 4-6), into the
 * context's ei/ej/ek like mhd_eijk_lfm (mhd_fetch, mhd_advance read them).
 * Binary64 contexts on the tensor-product grid (MHD_ERR_UNSUPPORTED
 * otherwise); the ghost layers must be filled, as for every routine here.
 * get_flux_stress_old (2131-2383): get_flux_stress's arithmetic, operation
 * for operation, on the Adams-Bashforth extrapolation rhoh, vxh, ... =
 * mhd_flux_stress(ctx, MHD_STATE_PREDICTED). */
int mhd_eijk_yee(mhd_ctx *ctx, int which_state);
int mhd_flux_stress_old(mhd_ctx *ctx);

/* the module variables tau and tau0 (module pars, mhd3d3.f90:33-248): the
 * step the update routines use and the previous one (both > 0); e.g. the
 * reference's `tau=tau0` of step_rk2 (3662) */
int mhd_set_tau(mhd_ctx *ctx, double tau, double tau0);
int mhd_get_tau(mhd_ctx *ctx, double *tau, double *tau0);

/* advance_predictor_ab (mhd3d3.f90:3821-3840) with its advance_previous and
 * the tau0=tau that follows it in step_ab (3616): the extrapolated state
 * x + tau/tau0/2 (x - x_p) of the 11 state arrays, then x_p = x.
 * Adams-Bashforth contexts only (MHD_ERR_UNSUPPORTED otherwise). */
int mhd_predict_ab(mhd_ctx *ctx);

/* the update of one stage from the face fluxes and edge EMFs the context
 * holds (mhd_flux_stress / mhd_eijk_lfm, or mhd_store): get_conserve
 * (1641-1650), advance_corrector1 (3842-3878), the stage combine and
 * pressure, advance_corrector2 (3880-3907), the combine and velocity,
 * advance_bijk (3909-3933) and the face combine -- the body of
 * step_rk2/step_rk3 (3665-3707, 3733-3785) or step_ab (3622-3627) between
 * the flux and the ghost fill, in place on the current state.
 *   stage 0: Adams-Bashforth corrector (stepop 1);
 *   stage 1..stepop: Runge-Kutta stage; stage 1 first takes the
 *            advance_previous copies (3794-3819) of the state it starts
 *            from (step_rk2 takes them just before stage 1, 3664).
 * Cell B and the CFL partials follow at mhd_get_bxyz, as in the
 * reference.  Only interior cells and owned faces are written: the
 * reference's whole-array stage combine also averages the ghost layers,
 * which the ghost fill that follows overwrites in both.  MHD_ERR_PARAM for a
 * stage the context's stepper lacks. */
int mhd_advance(mhd_ctx *ctx, int stage);

/* copy a device product to the host:
 *   which = MHD_FETCH_FLUX + q (q = 8*d + f, d = face direction x/y/z,
 *           f = rho, rvx, rvy, rvz, eng, stress_x, stress_y, stress_z),
 *           MHD_FETCH_EI/EJ/EK, MHD_FETCH_VF (cell array |v|+|B|/sqrt(rho)) */
enum { MHD_FETCH_FLUX = 0, MHD_FETCH_EI = 24, MHD_FETCH_EJ = 25, MHD_FETCH_EK = 26, MHD_FETCH_VF = 27 };
int mhd_fetch(mhd_ctx *ctx, int which, double *host_out);
/* the inverse of mhd_fetch for the face fluxes and edge EMFs (no
 * counterpart in the reference, whose flux arrays are module variables any
 * routine may write, mhd3d3.f90:273-280): the next mhd_advance uses them.
 * Mixed precision stores them as fp32.  MHD_ERR_PARAM for other products. */
int mhd_store(mhd_ctx *ctx, int which, const double *host_in);

/* reconstruct_3d (mhd3d3.f90:833-860) of a host array f(fn1,fn2,fn3) into
 * fl, fr (lnx,lny,lnz); dir 1-3 truncated, 4-6 untruncated transverse */
int mhd_reconstruct(mhd_ctx *ctx, int dir, const double *f, int fn1, int fn2, int fn3, int lnx, int lny,
                    int lnz, double *fl, double *fr);

/* get_v_face (mhd3d3.f90:2594-2615) elementwise on n host values, in the
 * arithmetic of the context's flux: binary64 (ROCm erfc/exp) or, in mixed
 * precision, fp32 (erfc by Abramowitz & Stegun 7.1.26 sharing the thermal
 * exponential).  v0p = erfc(-a v)/2, v0n = erfc(a v)/2, v1p = v v0p + b,
 * v1n = v v0n - b, a = sqrt(lam), b = exp(-lam v^2)/(2 a sqrt(pi)). */
int mhd_v_face(mhd_ctx *ctx, long n, const double *lam, const double *v, double *v0p, double *v0n, double *v1p,
               double *v1n);

/* mhd_reconstruct without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as mhd_reconstruct, except that f, fl
 * and fr are staged through one fixed device array per GPU of
 * MHD_NOALLOC_DOUBLES doubles (32 MiB, held until the process ends):
 * fn1*fn2*fn3 + 2*lnx*lny*lnz, each rounded up to a multiple of 32, must not
 * exceed it (MHD_ERR_PARAM otherwise).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (from several threads or
 * contexts) race and must be serialised by the caller. */
enum { MHD_NOALLOC_DOUBLES = 1 << 22 };
int mhd_reconstruct_noalloc(mhd_ctx *ctx, int dir, const double *f, int fn1, int fn2, int fn3, int lnx, int lny,
                            int lnz, double *fl, double *fr);

/* divergence monitor (no counterpart in the reference, which has no div B
 * diagnostic): *divb = max over the owned cells of |div B| of the current
 * face field, each cell's net face flux over its volume as the update forms
 * a flux divergence (metric included).  CT keeps it at round-off. */
int mhd_divb_max(mhd_ctx *ctx, double *divb);

/* diag_rho_neg / diag_overflow (mhd3d3.f90:3143-3160) plus p<=0:
 * *flags = 1 rho<0, 2 NaN interior rho, 4 interior p<=0 or NaN;
 * returns MHD_ERR_STATE when any flag is set */
int mhd_check_state(mhd_ctx *ctx, int *flags);

int mhd_synchronize(mhd_ctx *ctx);

/* device pointer of a state array (for zero-copy callers):
 * 0 rho 1 vx 2 vy 3 vz 4 p 5 bx 6 by 7 bz 8 bi 9 bj 10 bk */
void *mhd_device_pointer(mhd_ctx *ctx, int which);

/* kernel timing with HIP events on the context's stream (no counterpart in
 * the reference, which only prints dates per frame, mhd3d3.f90:3292-3295).
 * kernel classes: see MHD_K_*; *ms = summed duration, *launches = count */
enum { MHD_K_FLUX = 0, MHD_K_EMF = 1, MHD_K_UPDATE = 2, MHD_K_CT = 3, MHD_K_FILL = 4, MHD_K_BXYZ = 5,
       MHD_K_DT = 6, MHD_K_PREV = 7, MHD_K_SURR = 8, MHD_K_NCLASS = 9 };
/* (the learned-reconstruction flux's first pass counts as MHD_K_FLUX;
 * MHD_K_SURR covers its MLP pass over the listed faces, the PDM fallback
 * fix-up, the list reset, sample collection, training and standalone
 * inference) */
int mhd_profile(mhd_ctx *ctx, int enable);
int mhd_profile_read(mhd_ctx *ctx, int kclass, int *launches, double *ms);
/* time only the classes whose bit (1 << MHD_K_*) is set in mask (0: off);
 * each timed launch is bracketed by two event records, which cost the
 * stream some time: a caller timing a whole step brackets only the class it
 * needs */
int mhd_profile_classes(mhd_ctx *ctx, unsigned mask);

/* stage scheduling (no counterpart in the reference): on (the default for
 * blocks of at least 2^21 owned cells), a stage's edge EMF and CT run on a
 * second stream beside its face flux and cell update; off (the default for
 * smaller blocks, and for ensemble members), every kernel runs on the
 * context's stream in the reference's order (the per-kernel times then add
 * up to the step).  The results are the same bit for bit either way. */
int mhd_set_stage_overlap(mhd_ctx *ctx, int on);

/* kernel fusion (no counterpart in the reference): on, a mixed-precision
 * Adams-Bashforth step with plain PDM on the tensor-product or affine grid
 * forms the fp32 face fluxes and applies the binary64 corrector in one
 * kernel, so the face products never reach HBM; off (the default: on
 * MI355X the fused kernel is slightly slower, its register footprint
 * leaving two wavefronts per SIMD), the flux kernels store the face
 * products and the update kernel reads them back.  The two agree to the
 * rounding of the fp32 flux (the fused kernel's floating-point contraction
 * may round some face values differently in the last place). */
int mhd_set_fusion(mhd_ctx *ctx, int on);

/* ---- learned face reconstruction (no counterpart in the reference) ------
 * A net y = lin(x) + gate(x) * MLP(x), MLP 40 -> 32 -> 32 -> 16 (activation
 * z*sigmoid(1.702 z)); per variable lin is the fixed 4th-order interpolation
 * (-x0 + 7x1 + 7x2 - x3)/12 and gate the largest |second difference| of its
 * window.  It maps the normalised 5-point
 * stencils (cells j-2..j+2 of the face between j-1 and j) of the 8
 * primitives to the face's left/right states; it replaces PDM in the fp32
 * flux of a MHD_MIXED context created with surrogate = 1, with PDM as
 * fallback where the largest normalised second difference exceeds theta or
 * a density/pressure state is not positive, and with the linear part
 * alone where it is below 1e-3 (the gate keeps the net's correction within
 * 1e-3 of the window scale there).  Windows are centred on the
 * mean of the two cells next to the face and divided by max(range, 1e-3 x
 * the variable's physical scale); states are clamped to the two cells'
 * range.
 * Parameters: MHD_SUR_NPARAM fp32, packed W1[32][40] b1[32] W2[32][32]
 * b2[32] W3[16][32] b3[16]; inference uses bf16 copies (MFMA), fp32
 * accumulate.  Samples: 56 floats = 40 normalised inputs + 16 PDM targets. */
enum { MHD_SUR_NPARAM = 2896, MHD_SUR_SAMPLE = 56 };
int mhd_surrogate_set_weights(mhd_ctx *ctx, const float *w, int n);
int mhd_surrogate_get_weights(mhd_ctx *ctx, float *w, int n);
/* theta: smoothness threshold of the net (finite; theta < 0 sends every
 * face to the PDM fallback) */
int mhd_surrogate_configure(mhd_ctx *ctx, double theta, int count_fallbacks);
/* faces evaluated / PDM fallbacks since the last call (count_fallbacks on) */
int mhd_surrogate_stats(mhd_ctx *ctx, long *faces, long *fallbacks);
/* append face samples of the current state (faces with t % stride == phase,
 * phase cycling per call) whose roughness <= theta; capacity = set size */
int mhd_surrogate_collect(mhd_ctx *ctx, double theta, int stride, long capacity, long *collected);
int mhd_surrogate_load_samples(mhd_ctx *ctx, const float *samples, long n);
int mhd_surrogate_fetch_samples(mhd_ctx *ctx, float *out, long n);
/* Adam training on the collected set (GPU forward/backward kernels) */
int mhd_surrogate_train(mhd_ctx *ctx, int epochs, int batch, double lr, unsigned seed, double *mean_loss);
/* loss and gradient of the given weights on n host samples (gradient checks);
 * the context's own net and optimiser state are not touched */
int mhd_surrogate_grad(mhd_ctx *ctx, const float *w, const float *samples, int n, double *loss, float *grad);
/* bf16-MFMA forward of n host rows of 40 normalised inputs -> n rows of 16 */
int mhd_surrogate_infer(mhd_ctx *ctx, const float *x, long n, float *y);

#ifdef __cplusplus
}
#endif
#endif
