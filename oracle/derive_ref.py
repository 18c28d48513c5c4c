#!/usr/bin/env python3
"""Build the reference Fortran solver into oracle/_ref/ (test infrastructure only).

This recipe compiles ``mhd3d3.f90`` where it lies under /root/reference with the
image's ROCm ``flang`` and the image's MPICH (/opt/conda).  The reference is
built with ifort ``-r8`` (makefile:19): default REAL becomes 8 bytes while
DOUBLE PRECISION stays 8 bytes.  flang's ``-fdefault-real-8`` alone would also
promote DOUBLE PRECISION (hence ``dble()``) to REAL(16), which changes the
grid coordinates of init_geom (mhd3d3.f90:3395-3397) in the last bit, so the
recipe adds ``-fdefault-double-8`` to get the ifort semantics.  The reference source is never
copied into the repository: a patched copy is written to a temporary directory
outside the tree, compiled, and deleted; only the executables land in
oracle/_ref/ (git-ignored).

Patches applied to the temporary copy (none touches the numerics):
  * remove module ``output_h5`` (mhd3d3.f90:2811-3129) and its five call lines:
    the image's HDF5 Fortran module is gfortran-format and unusable by flang;
    frame output is not on the step path;
  * ``init_ic`` (mhd3d3.f90:3517) first looks for ``ic.bin`` (global interior
    arrays, column-major float64: rho,vx,vy,vz,p (nx0G,ny0G,nz0G), then
    bi (nx0G+1,..), bj, bk) and, if present, loads the rank's slab; otherwise
    the reference's own shock-tube initial state is used unchanged;
  * raw dumps (stream, little-endian) of the owned state after
    initialization and after every frame, and of the per-routine outputs
    (get_eijk_LFM, get_flux_stress, get_dt, and the alternative
    get_eijk_Yee whose own calls the reference comments out) evaluated on
    the initial state, when a file ``dump.on`` exists in the run directory;
  * the step loop (mhd3d3.f90:3283-3299) is timed with MPI_Wtime and the
    wall time printed by rank 0 as ``STEPLOOP_SECONDS <t>``;
  * the ``periodic`` variant additionally sets CART_BCs=.TRUE.
    (mhd3d3.f90:3204-3206), drops the PROC_NULL overrides (3229-3230,
    3242-3243, 3255-3256) and passes boundary code 2 (a no-op) to the fill
    calls (781-819), giving the periodic halo semantics of mhd3d2.f90:3061-3096.

Usage:  python3 oracle/derive_ref.py            (builds both variants)
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

REF = "/root/reference/mhd3d3.f90"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib"

EXTRA = r"""
!>>>>> oracle harness additions (not part of the reference numerics) >>>>>
subroutine read_ic_file
   use pars
   use fields
   implicit none
   real*8, allocatable :: g(:,:,:)
   integer :: u, i0, j0, k0
   i0 = xID*nx0
   j0 = yID*ny0
   k0 = zID*nz0
   open(newunit=u, file='ic.bin', access='stream', form='unformatted', status='old')
   allocate(g(nx0G,ny0G,nz0G))
   read(u) g
   rho(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0) = g(i0+1:i0+nx0,j0+1:j0+ny0,k0+1:k0+nz0)
   read(u) g
   vx(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0) = g(i0+1:i0+nx0,j0+1:j0+ny0,k0+1:k0+nz0)
   read(u) g
   vy(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0) = g(i0+1:i0+nx0,j0+1:j0+ny0,k0+1:k0+nz0)
   read(u) g
   vz(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0) = g(i0+1:i0+nx0,j0+1:j0+ny0,k0+1:k0+nz0)
   read(u) g
   p(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0) = g(i0+1:i0+nx0,j0+1:j0+ny0,k0+1:k0+nz0)
   deallocate(g)
   allocate(g(nx0G+1,ny0G,nz0G))
   read(u) g
   bi(ng+1:ng+nx0+1,ng+1:ng+ny0,ng+1:ng+nz0) = g(i0+1:i0+nx0+1,j0+1:j0+ny0,k0+1:k0+nz0)
   deallocate(g)
   allocate(g(nx0G,ny0G+1,nz0G))
   read(u) g
   bj(ng+1:ng+nx0,ng+1:ng+ny0+1,ng+1:ng+nz0) = g(i0+1:i0+nx0,j0+1:j0+ny0+1,k0+1:k0+nz0)
   deallocate(g)
   allocate(g(nx0G,ny0G,nz0G+1))
   read(u) g
   bk(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0+1) = g(i0+1:i0+nx0,j0+1:j0+ny0,k0+1:k0+nz0+1)
   deallocate(g)
   close(u)
end subroutine read_ic_file
!*********************************************
logical function dump_enabled()
   implicit none
   inquire(file='dump.on', exist=dump_enabled)
end function dump_enabled
!*********************************************
subroutine dump_state(tag)
   use pars
   use fields
   implicit none
   character(len=*) :: tag
   character(len=96) :: fname
   integer :: u
   logical, external :: dump_enabled
   if (.not. dump_enabled()) return
   write(fname,'(A,A,A,I0,A)') 'dump_', trim(tag), '_r', myrank, '.bin'
   open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
   write(u) xID, yID, zID, nx0, ny0, nz0
   write(u) time, tau
   write(u) rho(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0)
   write(u) vx(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0)
   write(u) vy(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0)
   write(u) vz(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0)
   write(u) p(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0)
   write(u) bi(ng+1:ng+nx0+1,ng+1:ng+ny0,ng+1:ng+nz0)
   write(u) bj(ng+1:ng+nx0,ng+1:ng+ny0+1,ng+1:ng+nz0)
   write(u) bk(ng+1:ng+nx0,ng+1:ng+ny0,ng+1:ng+nz0+1)
   close(u)
end subroutine dump_state
!*********************************************
subroutine dump_frame(n)
   implicit none
   integer :: n
   character(len=16) :: tag
   write(tag,'(A,I0)') 'f', n
   call dump_state(trim(tag))
end subroutine dump_frame
!*********************************************
subroutine dump_init_ops
!..per-routine outputs on the initial (filled) state
   use pars
   use fields
   use get
   implicit none
   character(len=96) :: fname
   integer :: u
   real*8 :: tsave, tcfl
   real*8, allocatable :: sei(:,:,:), sej(:,:,:), sek(:,:,:)
   logical, external :: dump_enabled
   if (.not. dump_enabled()) return
   tsave = tau
   call get_dt
   tcfl = tau
   tau = tsave
   call get_flux_stress(rho,vx,vy,vz,p,bx,by,bz)
   write(fname,'(A,I0,A)') 'dump_ops_r', myrank, '.bin'
   open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
   write(u) xID, yID, zID, nx0, ny0, nz0
   write(u) tcfl
   write(u) ei, ej, ek
   write(u) flux_rhox, flux_rvxx, flux_rvyx, flux_rvzx, flux_engx, stress_xx, stress_yx, stress_zx
   write(u) flux_rhoy, flux_rvxy, flux_rvyy, flux_rvzy, flux_engy, stress_xy, stress_yy, stress_zy
   write(u) flux_rhoz, flux_rvxz, flux_rvyz, flux_rvzz, flux_engz, stress_xz, stress_yz, stress_zz
   close(u)
!..the reference's alternative EMF get_eijk_Yee on the same state (its own
!..calls are commented out); ei, ej, ek are put back afterwards
   allocate(sei(size(ei,1),size(ei,2),size(ei,3)), sej(size(ej,1),size(ej,2),size(ej,3)), &
            sek(size(ek,1),size(ek,2),size(ek,3)))
   sei = ei
   sej = ej
   sek = ek
   call get_eijk_Yee(rho,vx,vy,vz,p,bx,by,bz)
   write(fname,'(A,I0,A)') 'dump_yee_r', myrank, '.bin'
   open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
   write(u) xID, yID, zID, nx0, ny0, nz0
   write(u) ei, ej, ek
   close(u)
   ei = sei
   ej = sej
   ek = sek
   deallocate(sei, sej, sek)
end subroutine dump_init_ops
!<<<<< oracle harness additions <<<<<
"""


def _replace_once(text, old, new):
    n = text.count(old)
    if n != 1:
        raise RuntimeError("patch anchor found %d times: %r" % (n, old[:80]))
    return text.replace(old, new)


def patched_source(periodic):
    src = open(REF).read()
    # 1. drop module output_h5 and its call lines
    a = src.index("module output_h5")
    a = src.rindex("\n", 0, a) + 1
    b = src.index("end module output_h5")
    b = src.index("\n", b) + 1
    src = src[:a] + src[b:]
    for line in ["   use output_h5\n", "   call init_output_h5\n",
                 "   call terminate_output_h5\n"]:
        src = src.replace(line, "")
    # the two outputxyz calls in the main program: the first (t=0) goes, the
    # per-frame one becomes a raw dump
    src = src.replace("   call outputxyz\n", "", 1)
    src = src.replace("      call outputxyz\n", "      call dump_frame(ntt)\n", 1)
    src = src.replace("   if(myrank.eq.0) print*,'hdf5 output terminated'\n", "")
    # 2. IC from file
    src = _replace_once(
        src,
        "subroutine init_ic\n   use pars\n   use fields\n   implicit none\n   integer :: i,j,k\n",
        "subroutine init_ic\n   use pars\n   use fields\n   implicit none\n   integer :: i,j,k\n"
        "   logical :: icfile\n   inquire(file='ic.bin', exist=icfile)\n"
        "   if (icfile) then\n      call read_ic_file\n      return\n   endif\n")
    # 3. dumps: initial state + per-routine ops after the diagnostic EMF call
    src = _replace_once(
        src,
        "   call get_eijk_LFM(rho,vx,vy,vz,p,bx,by,bz,bi,bj,bk)\n\n   if (stepop.eq.1) then\n",
        "   call get_eijk_LFM(rho,vx,vy,vz,p,bx,by,bz,bi,bj,bk)\n   call dump_init_ops\n"
        "   call dump_state('f0')\n\n   if (stepop.eq.1) then\n")
    # 4. step-loop timing
    src = _replace_once(src, "   character(10) :: rtime\n",
                        "   character(10) :: rtime\n   real*8 :: t_loop0, t_loop1\n")
    src = _replace_once(src, "   do ntt=1,nframes\n      do nt=1,nts\n",
                        "   call MPI_BARRIER(MPI_COMM_WORLD,ierr)\n   t_loop0=MPI_WTIME()\n"
                        "   do ntt=1,nframes\n      do nt=1,nts\n")
    src = _replace_once(src, "      call diag_overflow\n   enddo\n",
                        "      call diag_overflow\n   enddo\n"
                        "   call MPI_BARRIER(MPI_COMM_WORLD,ierr)\n   t_loop1=MPI_WTIME()\n"
                        "   if (myrank.eq.0) print '(A,ES24.16)', 'STEPLOOP_SECONDS ', t_loop1-t_loop0\n")
    if periodic:
        for ax in ("1", "2", "3"):
            src = re.sub(r"\n   CART_BCs\(%s\) = \.FALSE\.[^\n]*" % ax,
                         "\n   CART_BCs(%s) = .TRUE." % ax, src, count=1)
        for v in ("left", "right", "down", "up", "back", "front"):
            src = re.sub(r"\n   if \([xyz]ID==[^\n]*\) %s=MPI_PROC_NULL" % v, "\n", src, count=1)
        a = src.index("   subroutine fill_hydro_mpi")
        b = src.index("   end subroutine fill_mag_mpi")
        body = re.sub(r"(call fill[xyz]_(?:center|face)\([^)]*),0\)", r"\1,2)", src[a:b])
        src = src[:a] + body + src[b:]
    return src + EXTRA


def build(variant):
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "mhd3d3_ref_%s.x" % variant)
    tmp = tempfile.mkdtemp(prefix="mhdref_")
    try:
        f90 = os.path.join(tmp, "ref.f90")
        with open(f90, "w") as fh:
            fh.write(patched_source(variant == "periodic"))
        cmd = [FLANG, "-fdefault-real-8", "-fdefault-double-8", "-O3", "-I" + MPI_INC, f90, "-o", exe,
               "-L" + MPI_LIB, "-lmpifort", "-lmpi", "-Wl,-rpath," + MPI_LIB]
        r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr[-4000:])
            raise RuntimeError("flang failed for variant %s" % variant)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return exe


def available():
    return os.path.exists(REF) and os.path.exists(FLANG) and os.path.exists(MPI_LIB + "/libmpi.so")


if __name__ == "__main__":
    if not available():
        print("reference or toolchain absent; nothing built")
        sys.exit(0)
    for v in ("even", "periodic"):
        print("built", build(v))
