!-----------------------------------------------------------------------
! mhd_gpu_mod.f90 -- ISO_C_BINDING interface to libmhd_gpu.so
! (C declarations in include/mhd_gpu.h).
!
! The reference's per-step routines (step, get_dt, fill_hydro_mpi +
! fill_mag_mpi, get_bxyz, get_flux_stress, get_eijk_LFM; mhd3d3.f90) map
! one-to-one onto these bindings.  Arrays are passed as assumed-size
! real(c_double) dummies in the reference layout (column-major, x fastest,
! ghost-inclusive).
!-----------------------------------------------------------------------
module mhd_gpu
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: MHD_OK = 0, MHD_ERR_HIP = 1, MHD_ERR_PARAM = 2, &
                                MHD_ERR_STATE = 3, MHD_ERR_UNSUPPORTED = 4
   integer(c_int), parameter :: MHD_BC_EVEN = 0, MHD_BC_PERIODIC = 2
   integer(c_int), parameter :: MHD_FP64 = 0, MHD_MIXED = 1
   ! the state get_flux_stress / get_eijk_LFM read (their array arguments)
   integer(c_int), parameter :: MHD_STATE_CURRENT = 0, MHD_STATE_PREDICTED = 1

   type, bind(C) :: mhd_params
      integer(c_int) :: nx0, ny0, nz0, ng
      real(c_double) :: gam, cfl, pdmb, eps, alpha, beta
      integer(c_int) :: reconop, stepop
      integer(c_int) :: bc(3)
      integer(c_int) :: precision, surrogate, device
   end type mhd_params

   interface
      subroutine mhd_default_params(p) bind(C, name='mhd_default_params')
         import :: mhd_params
         type(mhd_params), intent(out) :: p
      end subroutine
      integer(c_int) function mhd_create(ctx, p) bind(C, name='mhd_create')
         import :: c_int, c_ptr, mhd_params
         type(c_ptr), intent(out) :: ctx
         type(mhd_params), intent(in) :: p
      end function
      subroutine mhd_destroy(ctx) bind(C, name='mhd_destroy')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine
      type(c_ptr) function mhd_error_string(code) bind(C, name='mhd_error_string')
         import :: c_ptr, c_int
         integer(c_int), value :: code
      end function
      integer(c_int) function mhd_set_geometry(ctx, dx, dy, dz) bind(C, name='mhd_set_geometry')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: dx(*), dy(*), dz(*)
      end function
      integer(c_int) function mhd_upload_state(ctx, rho, vx, vy, vz, p, bi, bj, bk) &
            bind(C, name='mhd_upload_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: rho(*), vx(*), vy(*), vz(*), p(*), bi(*), bj(*), bk(*)
      end function
      integer(c_int) function mhd_download_state(ctx, rho, vx, vy, vz, p, bi, bj, bk, bx, by, bz) &
            bind(C, name='mhd_download_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: rho(*), vx(*), vy(*), vz(*), p(*), bi(*), bj(*), bk(*), &
                                        bx(*), by(*), bz(*)
      end function
      integer(c_int) function mhd_initialize(ctx, tau) bind(C, name='mhd_initialize')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: tau
      end function
      integer(c_int) function mhd_step(ctx, nsteps, tau, tau_cfl) bind(C, name='mhd_step')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsteps
         real(c_double), intent(inout) :: tau
         real(c_double), intent(out) :: tau_cfl
      end function
      integer(c_int) function mhd_get_dt(ctx, tau_cfl) bind(C, name='mhd_get_dt')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: tau_cfl
      end function
      integer(c_int) function mhd_fill_ghosts(ctx) bind(C, name='mhd_fill_ghosts')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mhd_get_bxyz(ctx) bind(C, name='mhd_get_bxyz')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mhd_flux_stress(ctx, which_state) bind(C, name='mhd_flux_stress')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: which_state
      end function
      integer(c_int) function mhd_eijk_lfm(ctx, which_state) bind(C, name='mhd_eijk_lfm')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: which_state
      end function
      ! the reference's dead alternatives get_eijk_Yee and get_flux_stress_old
      integer(c_int) function mhd_eijk_yee(ctx, which_state) bind(C, name='mhd_eijk_yee')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: which_state
      end function
      integer(c_int) function mhd_flux_stress_old(ctx) bind(C, name='mhd_flux_stress_old')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      ! the module variables tau, tau0
      integer(c_int) function mhd_set_tau(ctx, tau, tau0) bind(C, name='mhd_set_tau')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: tau, tau0
      end function
      integer(c_int) function mhd_get_tau(ctx, tau, tau0) bind(C, name='mhd_get_tau')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: tau, tau0
      end function
      ! advance_predictor_ab (+ advance_previous, tau0=tau)
      integer(c_int) function mhd_predict_ab(ctx) bind(C, name='mhd_predict_ab')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      ! advance_corrector1/2 + stage combine + advance_bijk (stage 0: AB)
      integer(c_int) function mhd_advance(ctx, stage) bind(C, name='mhd_advance')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: stage
      end function
      integer(c_int) function mhd_store(ctx, which, a) bind(C, name='mhd_store')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: which
         real(c_double), intent(in) :: a(*)
      end function
      integer(c_int) function mhd_fetch(ctx, which, out) bind(C, name='mhd_fetch')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: which
         real(c_double), intent(out) :: out(*)
      end function
      integer(c_int) function mhd_check_state(ctx, flags) bind(C, name='mhd_check_state')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: flags
      end function
      ! max |div B| over the owned cells (no reference counterpart)
      integer(c_int) function mhd_divb_max(ctx, divb) bind(C, name='mhd_divb_max')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: divb
      end function
      integer(c_int) function mhd_synchronize(ctx) bind(C, name='mhd_synchronize')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      ! affine metric x_phys = A x_comp, A row-major (no reference counterpart)
      integer(c_int) function mhd_set_metric(ctx, a) bind(C, name='mhd_set_metric')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: a(9)
      end function
      ! general curvilinear metric from the (nx+1)(ny+1)(nz+1) node
      ! coordinates, column-major, ghosts included (no reference counterpart)
      integer(c_int) function mhd_set_metric_nodes(ctx, x, y, z) bind(C, name='mhd_set_metric_nodes')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: x(*), y(*), z(*)
      end function
      ! batched ensemble of Runge-Kutta contexts, one graph launch per step
      ! (no reference counterpart)
      integer(c_int) function mhd_ensemble_create(ens, members, n, tau) bind(C, name='mhd_ensemble_create')
         import :: c_int, c_ptr, c_double
         type(c_ptr), intent(out) :: ens
         type(c_ptr), intent(in) :: members(*)
         integer(c_int), value :: n
         real(c_double), intent(in) :: tau(*)
      end function
      integer(c_int) function mhd_ensemble_step(ens, nsteps) bind(C, name='mhd_ensemble_step')
         import :: c_int, c_ptr
         type(c_ptr), value :: ens
         integer(c_int), value :: nsteps
      end function
      subroutine mhd_ensemble_destroy(ens) bind(C, name='mhd_ensemble_destroy')
         import :: c_ptr
         type(c_ptr), value :: ens
      end subroutine
      ! learned reconstruction (no reference counterpart)
      integer(c_int) function mhd_surrogate_set_weights(ctx, w, n) bind(C, name='mhd_surrogate_set_weights')
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: ctx
         real(c_float), intent(in) :: w(*)
         integer(c_int), value :: n
      end function
      integer(c_int) function mhd_surrogate_configure(ctx, theta, count) bind(C, name='mhd_surrogate_configure')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: theta
         integer(c_int), value :: count
      end function
      integer(c_int) function mhd_surrogate_stats(ctx, faces, fallbacks) bind(C, name='mhd_surrogate_stats')
         import :: c_int, c_ptr, c_long
         type(c_ptr), value :: ctx
         integer(c_long), intent(out) :: faces, fallbacks
      end function
   end interface

   integer(c_int), parameter :: MHD_SUR_NPARAM = 2896

contains

   ! stop with the library's message on a failed call (the reference stops
   ! with a message on every fatal condition, mhd3d3.f90:3139-3155)
   subroutine mhd_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (rc == MHD_OK) return
      call c_f_pointer(mhd_error_string(rc), msg, [128])
      n = 0
      do while (n < 128)
         if (msg(n+1) == c_null_char) exit
         n = n + 1
      enddo
      print '(A,A,A,I0,A,128A1)', 'mhd_gpu: ', where, ' failed (code ', rc, '): ', msg(1:n)
      stop 1
   end subroutine mhd_check

end module mhd_gpu
