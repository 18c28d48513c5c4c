!-----------------------------------------------------------------------
! mhd3d_gpu.f90 -- driver of the MI355X ideal-MHD step.
!
! Reads the reference's namelist deck (groups &dom &grid &comset &list
! &free, mhd3d3.f90:152-156) plus an optional &gpu group, builds the
! tensor-product geometry as init_geom does (mhd3d3.f90:3379-3515), sets
! the initial state (the reference's shock tube, mhd3d3.f90:3517-3569, or
! an initial-state file), and runs the frame loop of the reference main
! program (mhd3d3.f90:3283-3299) with every step on the GPU through the
! C-ABI.  The whole domain is one block on one GPU: xprocs/yprocs/zprocs are
! read and ignored.
!
! usage:  mhd3d_gpu.x [deck]          (default deck: mhd3d3.in)
!
! &gpu  bc_x, bc_y, bc_z   'even' (reference, default) or 'periodic'
!       precision          'fp64' (default) or 'mixed'
!       surrogate          0 (default) or 1: learned face reconstruction
!                          (precision 'mixed', reconop 1), weights from
!       weights            a file of 2896 float32 (written by the Python
!                          package's surrogate.save_weights)
!       theta              its smoothness threshold (default 0.25)
!       metric             9 reals, row-major A of an affine metric
!                          x_phys = A x_comp (e.g. a sheared box); default
!                          the identity, which keeps the reference's grid
!       device             HIP device ordinal
!       icfile             initial-state file (global owned arrays, float64,
!                          column-major: rho vx vy vz p (nx0G,ny0G,nz0G),
!                          bi (nx0G+1,..), bj, bk); empty = shock tube
!       outdir             output directory (empty = no output): the
!                          reference's outputxyz datasets and stime.out
!                          (see write_xyz) and, with
!       frames             .true., also one raw file frame_<n>.bin per frame
!                          with the face fields (see write_frame)
!-----------------------------------------------------------------------
program mhd3d_gpu
   use, intrinsic :: iso_c_binding
   use mhd_gpu
   implicit none

   ! deck variables (module pars of the reference, mhd3d3.f90:66-91)
   real(8) :: xmin, xmax, ymin, ymax, zmin, zmax
   integer :: nx0G, ny0G, nz0G, xprocs, yprocs, zprocs
   integer :: nord, reconop, stepop, outop
   real(8) :: tau
   integer :: nts, nframes, nrst, nfdump
   real(8) :: cfl, pdmb, eps, alpha, beta
   character(len=16) :: bc_x, bc_y, bc_z, precision
   integer :: surrogate, device
   character(len=512) :: icfile, outdir, weights
   real(8) :: theta, metric(9)
   logical :: frames

   namelist /dom/ xmin, xmax, ymin, ymax, zmin, zmax
   namelist /grid/ nx0G, ny0G, nz0G, xprocs, yprocs, zprocs
   namelist /comset/ nord, reconop, stepop, outop
   namelist /list/ nts, tau, nframes, nrst, nfdump
   namelist /free/ cfl, pdmb, eps, alpha, beta
   namelist /gpu/ bc_x, bc_y, bc_z, precision, surrogate, device, icfile, outdir, weights, theta, metric, frames

   real(8), allocatable :: rho(:,:,:), vx(:,:,:), vy(:,:,:), vz(:,:,:), p(:,:,:)
   real(8), allocatable :: bx(:,:,:), by(:,:,:), bz(:,:,:)
   real(8), allocatable :: bi(:,:,:), bj(:,:,:), bk(:,:,:)
   real(8), allocatable :: xx0(:), yy0(:), zz0(:), dx(:), dy(:), dz(:)
   type(mhd_params) :: prm
   type(c_ptr) :: ctx
   character(len=512) :: deck
   integer :: ng, nx, ny, nz, ntt, nt, u, ios, flags
   integer(8) :: c0, c1, crate
   real(8) :: time, tcfl, wall
   integer(c_int) :: rc

   deck = 'mhd3d3.in'
   if (command_argument_count() >= 1) call get_command_argument(1, deck)

   ! defaults of the &gpu group
   bc_x = 'even'; bc_y = 'even'; bc_z = 'even'; precision = 'fp64'
   surrogate = 0; device = 0; icfile = ''; outdir = ''; weights = ''; theta = 0.25d0
   metric = [1.d0, 0.d0, 0.d0, 0.d0, 1.d0, 0.d0, 0.d0, 0.d0, 1.d0]
   frames = .false.
   xprocs = 1; yprocs = 1; zprocs = 1; outop = 0; nrst = 0; nfdump = 0

   print*, 'reading input file ', trim(deck)
   open(newunit=u, file=trim(deck), status='old')
   read(u, dom)
   read(u, grid)
   read(u, comset)
   read(u, list)
   read(u, free)
   read(u, gpu, iostat=ios)
   close(u)

   ng = nord/2
   nx = nx0G + nord
   ny = ny0G + nord
   nz = nz0G + nord

   ! geometry (init_geom): per-index metric, ghost faces mirrored
   allocate(xx0(nx+1), yy0(ny+1), zz0(nz+1), dx(nx), dy(ny), dz(nz))
   call axis_faces(xmin, xmax, nx0G, ng, xx0)
   call axis_faces(ymin, ymax, ny0G, ng, yy0)
   call axis_faces(zmin, zmax, nz0G, ng, zz0)
   dx = xx0(2:nx+1) - xx0(1:nx)
   dy = yy0(2:ny+1) - yy0(1:ny)
   dz = zz0(2:nz+1) - zz0(1:nz)
   if (minval(dx) < 0 .or. minval(dy) < 0 .or. minval(dz) < 0) stop 'Wrong grid! Simulation terminated!'

   allocate(rho(nx,ny,nz), vx(nx,ny,nz), vy(nx,ny,nz), vz(nx,ny,nz), p(nx,ny,nz))
   allocate(bx(nx,ny,nz), by(nx,ny,nz), bz(nx,ny,nz))
   allocate(bi(nx+1,ny,nz), bj(nx,ny+1,nz), bk(nx,ny,nz+1))
   rho = 0.; vx = 0.; vy = 0.; vz = 0.; p = 0.; bi = 0.; bj = 0.; bk = 0.
   if (len_trim(icfile) > 0) then
      call read_ic(trim(icfile))
   else
      call shock_tube
   endif

   call mhd_default_params(prm)
   prm%nx0 = nx0G; prm%ny0 = ny0G; prm%nz0 = nz0G; prm%ng = ng
   prm%gam = 5.d0/3.d0
   prm%cfl = cfl; prm%pdmb = pdmb; prm%eps = eps; prm%alpha = alpha; prm%beta = beta
   prm%reconop = reconop; prm%stepop = stepop
   prm%bc(1) = bc_code(bc_x); prm%bc(2) = bc_code(bc_y); prm%bc(3) = bc_code(bc_z)
   prm%precision = MHD_FP64
   if (trim(precision) == 'mixed') prm%precision = MHD_MIXED
   prm%surrogate = surrogate
   prm%device = device

   call mhd_check(mhd_create(ctx, prm), 'mhd_create')
   call mhd_check(mhd_set_geometry(ctx, dx, dy, dz), 'mhd_set_geometry')
   if (any(metric /= [1.d0, 0.d0, 0.d0, 0.d0, 1.d0, 0.d0, 0.d0, 0.d0, 1.d0])) &
      call mhd_check(mhd_set_metric(ctx, metric), 'mhd_set_metric')
   if (surrogate /= 0) call load_weights(trim(weights))
   call mhd_check(mhd_upload_state(ctx, rho, vx, vy, vz, p, bi, bj, bk), 'mhd_upload_state')
   call mhd_check(mhd_initialize(ctx, tau), 'mhd_initialize')
   print*, 'initialization finished!'

   time = 0.
   if (len_trim(outdir) > 0) call output(0)

   call system_clock(c0, crate)
   do ntt = 1, nframes
      if (stepop == 1) then
         ! Adams-Bashforth: tau changes every step, time advances by the
         ! previous tau before each step (mhd3d3.f90:3285)
         do nt = 1, nts
            time = time + tau
            call mhd_check(mhd_step(ctx, 1_c_int, tau, tcfl), 'mhd_step')
         enddo
      else
         time = time + nts*tau
         call mhd_check(mhd_step(ctx, int(nts, c_int), tau, tcfl), 'mhd_step')
         if (tcfl < tau) print*, 0, 'Attention! tau_cfl<', tau
      endif
      rc = mhd_check_state(ctx, flags)
      if (rc == MHD_ERR_STATE) then
         if (iand(flags, 1) /= 0) stop 'Density becomes negative! Simulation terminated!'
         if (iand(flags, 2) /= 0) stop 'NaN value found. Simulation terminated!'
         stop 'Non-positive pressure found. Simulation terminated!'
      endif
      call mhd_check(rc, 'mhd_check_state')
      if (len_trim(outdir) > 0) call output(ntt)
      print '(A,I0,A,ES14.6)', ' output frame number ', ntt+1, ' time=', time
   enddo
   call mhd_check(mhd_synchronize(ctx), 'mhd_synchronize')
   call system_clock(c1)
   wall = real(c1 - c0, 8)/real(crate, 8)
   print '(A,ES24.16)', 'STEPLOOP_SECONDS ', wall
   print '(A,ES14.6)', 'ZONE_UPDATES_PER_SECOND ', &
      real(nx0G,8)*real(ny0G,8)*real(nz0G,8)*real(nts,8)*real(nframes,8)/wall
   call mhd_destroy(ctx)

contains

   ! trained weights of the learned reconstruction: MHD_SUR_NPARAM float32
   subroutine load_weights(path)
      character(len=*), intent(in) :: path
      real(c_float) :: w(MHD_SUR_NPARAM)
      integer :: uw
      if (len(path) == 0) stop 'surrogate = 1 needs weights = <file>'
      open(newunit=uw, file=path, access='stream', form='unformatted', status='old')
      read(uw) w
      close(uw)
      call mhd_check(mhd_surrogate_set_weights(ctx, w, MHD_SUR_NPARAM), 'mhd_surrogate_set_weights')
      call mhd_check(mhd_surrogate_configure(ctx, theta, 0_c_int), 'mhd_surrogate_configure')
   end subroutine load_weights

   integer(c_int) function bc_code(s)
      character(len=*), intent(in) :: s
      bc_code = MHD_BC_EVEN
      if (trim(s) == 'periodic') bc_code = MHD_BC_PERIODIC
   end function bc_code

   ! ghost-inclusive face coordinates along one axis (mhd3d3.f90:3391-3424)
   subroutine axis_faces(lo, hi, n0, ngh, xx)
      real(8), intent(in) :: lo, hi
      integer, intent(in) :: n0, ngh
      real(8), intent(out) :: xx(:)
      real(8), allocatable :: x0(:)
      real(8) :: d0
      integer :: i
      allocate(x0(n0+1))
      d0 = (hi - lo)/n0
      do i = 1, n0+1
         x0(i) = real(i-1, 8)*d0 + lo
      enddo
      xx(ngh+1:ngh+n0+1) = x0
      do i = 1, ngh
         xx(i) = 2.d0*x0(1) - x0(ngh+2-i)
         xx(n0+2*ngh+2-i) = 2.d0*x0(n0+1) - x0(n0-ngh+i)
      enddo
      deallocate(x0)
   end subroutine axis_faces

   ! the reference's initial state: shock tube along z (mhd3d3.f90:3517-3569)
   subroutine shock_tube
      integer :: i, j, k
      real(8) :: zc, zi
      rho = 1.d0; p = 1.d0; vx = 0.d0; vy = 0.d0; vz = 0.d0
      bi = -1.d0; bj = 0.d0; bk = .75d0
      do k = 1, nz
         zc = .125d0*(zz0(k) + zz0(k+1) + zz0(k+1) + zz0(k) + zz0(k) + zz0(k+1) + zz0(k+1) + zz0(k))
         zi = .25d0*(zz0(k) + zz0(k) + zz0(k+1) + zz0(k+1))
         do j = 1, ny
            do i = 1, nx
               if (zc >= 0.d0) then
                  rho(i,j,k) = .125d0
                  p(i,j,k) = .1d0
               endif
            enddo
            do i = 1, nx+1
               if (zi < 0.d0) bi(i,j,k) = 1.d0
            enddo
         enddo
      enddo
   end subroutine shock_tube

   subroutine read_ic(fname)
      character(len=*), intent(in) :: fname
      integer :: v
      real(8), allocatable :: g(:,:,:)
      open(newunit=v, file=fname, access='stream', form='unformatted', status='old')
      allocate(g(nx0G,ny0G,nz0G))
      read(v) g; rho(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G) = g
      read(v) g; vx(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G) = g
      read(v) g; vy(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G) = g
      read(v) g; vz(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G) = g
      read(v) g; p(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G) = g
      deallocate(g)
      allocate(g(nx0G+1,ny0G,nz0G)); read(v) g
      bi(ng+1:ng+nx0G+1, ng+1:ng+ny0G, ng+1:ng+nz0G) = g; deallocate(g)
      allocate(g(nx0G,ny0G+1,nz0G)); read(v) g
      bj(ng+1:ng+nx0G, ng+1:ng+ny0G+1, ng+1:ng+nz0G) = g; deallocate(g)
      allocate(g(nx0G,ny0G,nz0G+1)); read(v) g
      bk(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G+1) = g; deallocate(g)
      close(v)
   end subroutine read_ic

   ! frame output after frame n: the state comes back from the device once
   subroutine output(n)
      integer, intent(in) :: n
      call mhd_check(mhd_download_state(ctx, rho, vx, vy, vz, p, bi, bj, bk, bx, by, bz), &
                     'mhd_download_state')
      call write_xyz(n)
      if (frames) call write_frame(n)
   end subroutine output

   ! outputxyz (mhd3d3.f90:3091-3109) into the datasets createh5_3D makes
   ! (2966-3000): per variable rho vx vy vz p bx by bz one dataset of shape
   ! (nx0G, ny0G, nz0G, nframes+1), float64, column-major, frame n the n-th
   ! 3-D slab -- here a file <outdir>/<var>.xyz holding the dataset's bytes
   ! (little-endian, no header: the image has no HDF5 library).  Plus
   ! output_simt (3112-3127): stime.out, Fortran sequential unformatted, one
   ! record holding `time` per frame, frame 0 creating the file.
   subroutine write_xyz(n)
      integer, intent(in) :: n
      character(len=3), parameter :: names(8) = ['rho', 'vx ', 'vy ', 'vz ', 'p  ', 'bx ', 'by ', 'bz ']
      character(len=600) :: fname
      integer :: v, q, i0, i1, j0, j1, k0, k1
      integer(8) :: pos
      i0 = ng + 1; i1 = ng + nx0G; j0 = ng + 1; j1 = ng + ny0G; k0 = ng + 1; k1 = ng + nz0G
      pos = int(n, 8) * 8_8 * int(nx0G, 8) * int(ny0G, 8) * int(nz0G, 8) + 1_8
      do q = 1, 8
         fname = trim(outdir)//'/'//trim(names(q))//'.xyz'
         if (n == 0) then
            open(newunit=v, file=trim(fname), access='stream', form='unformatted', status='replace')
         else
            open(newunit=v, file=trim(fname), access='stream', form='unformatted', status='old')
         endif
         select case (q)
         case (1); write(v, pos=pos) rho(i0:i1, j0:j1, k0:k1)
         case (2); write(v, pos=pos) vx(i0:i1, j0:j1, k0:k1)
         case (3); write(v, pos=pos) vy(i0:i1, j0:j1, k0:k1)
         case (4); write(v, pos=pos) vz(i0:i1, j0:j1, k0:k1)
         case (5); write(v, pos=pos) p(i0:i1, j0:j1, k0:k1)
         case (6); write(v, pos=pos) bx(i0:i1, j0:j1, k0:k1)
         case (7); write(v, pos=pos) by(i0:i1, j0:j1, k0:k1)
         case (8); write(v, pos=pos) bz(i0:i1, j0:j1, k0:k1)
         end select
         close(v)
      enddo
      fname = trim(outdir)//'/stime.out'
      if (n == 0) then
         open(newunit=v, file=trim(fname), status='replace', form='unformatted')
      else
         open(newunit=v, file=trim(fname), status='old', form='unformatted', position='append')
      endif
      write(v) time
      close(v)
   end subroutine write_xyz

   ! owned state of one frame (stream, little-endian): six int32 (block
   ! coordinates 0 0 0, nx0G ny0G nz0G), time, tau, then rho vx vy vz p bi
   ! bj bk without ghosts -- the per-frame variables of outputxyz
   ! (mhd3d3.f90:3091-3109) in raw form
   subroutine write_frame(n)
      integer, intent(in) :: n
      character(len=600) :: fname
      integer :: v
      write(fname, '(A,A,I4.4,A)') trim(outdir), '/frame_', n, '.bin'
      open(newunit=v, file=trim(fname), access='stream', form='unformatted', status='replace')
      write(v) 0, 0, 0, nx0G, ny0G, nz0G
      write(v) time, tau
      write(v) rho(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G)
      write(v) vx(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G)
      write(v) vy(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G)
      write(v) vz(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G)
      write(v) p(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G)
      write(v) bi(ng+1:ng+nx0G+1, ng+1:ng+ny0G, ng+1:ng+nz0G)
      write(v) bj(ng+1:ng+nx0G, ng+1:ng+ny0G+1, ng+1:ng+nz0G)
      write(v) bk(ng+1:ng+nx0G, ng+1:ng+ny0G, ng+1:ng+nz0G+1)
      close(v)
   end subroutine write_frame

end program mhd3d_gpu
