/*
 * mhd_oracle.h -- CPU restatement of the reference ideal-MHD step
 * (mhd3d3.f90), used ONLY as the checker by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  It is test infrastructure, never
 * part of the product path.
 *
 * Layout: Fortran column-major, x fastest, 1-based in the formulas.
 *   cell arrays  (nx,ny,nz)       nx = nx0 + 2*ng
 *   bi (nx+1,ny,nz)  bj (nx,ny+1,nz)  bk (nx,ny,nz+1)
 *   ei (nx0,ny0+1,nz0+1)  ej (nx0+1,ny0,nz0+1)  ek (nx0+1,ny0+1,nz0)
 *   flux/stress along x (nx0+1,ny0,nz0), along y (nx0,ny0+1,nz0),
 *   along z (nx0,ny0,nz0+1)          (mhd3d3.f90:251-317)
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx0, ny0, nz0, ng;
    double gam, cfl, pdmb, eps, alpha, beta;
    int reconop;   /* 1 PDM, 2 TVD, 3 WENO5, 4 MPWENO5 */
    int stepop;    /* 1 AB2, 2 RK2, 3 RK3 */
    int bc[3];     /* per axis: 0 even reflection, 2 periodic */
} orc_params;

typedef struct orc_state orc_state;

/* flux array order used by orc_get_ops (per direction d = x,y,z):
 * rho, rvx, rvy, rvz, eng, stress_x, stress_y, stress_z  -> index 8*d + q */
enum { ORC_NFLUX = 24 };

orc_state *orc_create(const orc_params *p);
void orc_destroy(orc_state *s);
void orc_set_geometry(orc_state *s, const double *dx, const double *dy, const double *dz);
void orc_set_state(orc_state *s, const double *rho, const double *vx, const double *vy,
                   const double *vz, const double *p, const double *bi, const double *bj,
                   const double *bk);
void orc_get_state(const orc_state *s, double *rho, double *vx, double *vy, double *vz,
                   double *p, double *bi, double *bj, double *bk, double *bx, double *by,
                   double *bz);
/* initialization (mhd3d3.f90:3327-3377): fills, get_bxyz, AB2 set-up. */
void orc_initialize(orc_state *s, double *tau);
void orc_fill(orc_state *s);                 /* fill_hydro_mpi + fill_mag_mpi */
void orc_bxyz(orc_state *s);                 /* get_bxyz */
double orc_get_dt(orc_state *s);             /* get_dt (min over the ghost-inclusive block) */
void orc_flux_stress(orc_state *s);          /* get_flux_stress on the current state */
void orc_eijk(orc_state *s);                 /* get_eijk_LFM on the current state */
void orc_get_ops(const orc_state *s, double *ei, double *ej, double *ek, double *flux_all);
/* per-routine pieces of the time stepper (mhd3d3.f90:3602-3933):
 * the module variables tau/tau0; advance_predictor_ab with its
 * advance_previous and tau0=tau (returns 1 unless stepop is 1); the flux and
 * EMF of the current (0) or extrapolated (1) state; get_conserve plus the
 * update of stage 0 (Adams-Bashforth) or 1..stepop (RK; stage 1 first takes
 * advance_previous) from the held fluxes and EMFs (returns 1 for a stage the
 * stepper lacks); and writing those fluxes and EMFs (orc_get_ops layout). */
void orc_set_tau(orc_state *s, double tau, double tau0);
void orc_get_tau(const orc_state *s, double *tau, double *tau0);
int orc_predict_ab(orc_state *s);
void orc_flux_stress_state(orc_state *s, int which);
void orc_eijk_state(orc_state *s, int which);
/* get_eijk_Yee (mhd3d3.f90:1654-1752) of the current (0) or extrapolated (1) state */
void orc_eijk_yee_state(orc_state *s, int which);
int orc_advance(orc_state *s, int stage);
void orc_set_ops(orc_state *s, const double *ei, const double *ej, const double *ek, const double *flux_all);
/* nsteps calls of `step` (mhd3d3.f90:3585); *tau in/out as the deck tau / tau0. */
int orc_step(orc_state *s, int nsteps, double *tau, double *tau_cfl);

/* reconstruct_3d (mhd3d3.f90:833) on a standalone array f(fn1,fn2,fn3). */
void orc_reconstruct(const orc_params *p, const double *f, int fn1, int fn2, int fn3,
                     int lnx, int lny, int lnz, int dir, double *fl, double *fr);
/* get_v_face (mhd3d3.f90:2594) elementwise. */
void orc_v_face(int n, const double *lam, const double *v, double *v0p, double *v0n,
                double *v1p, double *v1n);
/* learned reconstruction (MLP 40-32-32-16, activation z*sigmoid(1.702 z)):
 * forward (h_round: bf16-round hidden activations) and loss/gradient */
void orc_mlp_forward(const float *w, const float *x, long n, int h_round, float *y);
double orc_mlp_loss_grad(const float *w, const float *samples, long n, double *grad);
/* use the net for the face states of the flux (PDM fallback; reconop must
 * be 1); w = NULL switches it off.  bf16_round: round weights, inputs and
 * hidden activations to bf16 like the GPU's MFMA inference. */
void orc_set_surrogate(orc_state *s, const float *w, double theta, int bf16_round);

/* set the number of OpenMP threads used by the oracle (0 = runtime default). */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
