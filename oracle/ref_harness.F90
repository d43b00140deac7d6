! ref_harness.F90 -- bind(C) entry points around the reference's own
! hot-path routines (hydro/ppmsolver.F90, hydro/order2interp.F90,
! hydro/riemannsolver.F90, hydro/riemann.F90, hydro/integrate.F90), compiled
! together with them from the reference tree by oracle/Makefile into
! oracle/_ref/libhydropad_ref.so.
!
! TEST INFRASTRUCTURE ONLY: used to pin the C restatement (hp_oracle.c) and
! to write the frozen vectors under tests/golden/, and as the CPU baseline
! of bench.py.  Nothing here is part of the product.
module hp_ref_harness
  use iso_c_binding
  implicit none
contains

  ! Set the 19 module scalars of hydropad/scalar.F90:98-99 from an array in
  ! hp_params member order (include/hydropad_gpu.h).
  subroutine ref_set_params(prm) bind(C, name="ref_set_params")
    use scalar
    real(c_double), intent(in) :: prm(19)
    gamma = prm(1); gamma1 = prm(2); rgamma = prm(3); rgamma1 = prm(4)
    m = prm(5); rm = prm(6); gf = prm(7)
    dt = prm(8); dx = prm(9); rdx = prm(10); rdtath = prm(11)
    at = prm(12); dat = prm(13); rat = prm(14)
    flatvalue = prm(15); mingradflat = prm(16); eta1 = prm(17); eta2 = prm(18)
    dmax = prm(19)
    mype = 1          ! not rank 0: silences ppmsolver's per-sweep prints
    call ref_mpi_init()
  end subroutine ref_set_params

  ! ppmsolver times itself with MPI_WTIME (hydro/ppmsolver.F90:72,464), which
  ! MPICH only serves after MPI_Init: start a singleton MPI once.
  subroutine ref_mpi_init()
    include 'mpif.h'
    logical :: started
    integer :: ierr0
    call MPI_Initialized(started, ierr0)
    if (.not. started) call MPI_Init(ierr0)
  end subroutine ref_mpi_init

  ! One STENCIL sweep of the reference, in place.  gforce is (3,nx,ny,nz).
  subroutine ref_ppmsolver(p, rho, vx, vy, vz, gforce, nx, ny, nz, nbound, ndir) &
       bind(C, name="ref_ppmsolver")
    integer(c_int), value :: nx, ny, nz, nbound, ndir
    real(c_double) :: p(nx, ny, nz), rho(nx, ny, nz), vx(nx, ny, nz)
    real(c_double) :: vy(nx, ny, nz), vz(nx, ny, nz), gforce(3, nx, ny, nz)
    integer :: inx, iny, inz, inb, idir
    external :: ppmsolver
    inx = nx; iny = ny; inz = nz; inb = nbound; idir = ndir
    ! cho3d/nes3d are unused under STENCIL (hydropad/alloc_arrays.F90:61-63)
    call ppmsolver(p, rho, vx, vy, vz, p, p, gforce, inx, iny, inz, inb, idir)
  end subroutine ref_ppmsolver

  ! order2interp on one 6-cell stencil; out(10) = pleft..vzright.
  subroutine ref_order2interp(pres, rho, vx, vy, vz, g, c, out) bind(C, name="ref_order2interp")
    real(c_double), intent(in) :: pres(6), rho(6), vx(6), vy(6), vz(6), g(6), c(6)
    real(c_double), intent(out) :: out(10)
    real(c_double) :: a1(6), a2(6), a3(6), a4(6), a5(6), a6(6), a7(6)
    integer :: ngrid, nmin, nmax, nb
    external :: order2interp
    a1 = pres; a2 = rho; a3 = vx; a4 = vy; a5 = vz; a6 = g; a7 = c
    ngrid = 6; nmin = 3; nmax = 4; nb = 3
    call order2interp(ngrid, nmin, nmax, nb, a1, a2, a3, a4, a5, a6, a7, &
         out(1), out(2), out(3), out(4), out(5), out(6), out(7), out(8), out(9), out(10))
  end subroutine ref_order2interp

  ! riemannsolver on one interface; in(10) as ref_order2interp's out,
  ! out(5) = pm, rhom, vxm, vym, vzm.
  subroutine ref_riemannsolver(rhobar, vxbar, in, out) bind(C, name="ref_riemannsolver")
    real(c_double), intent(in) :: rhobar(6), vxbar(6), in(10)
    real(c_double), intent(out) :: out(5)
    real(c_double) :: rb(6), vb(6), s(10)
    integer :: ngrid
    external :: riemannsolver
    rb = rhobar; vb = vxbar; s = in; ngrid = 6
    call riemannsolver(ngrid, rb, vb, s(1), s(2), s(3), s(4), s(5), s(6), s(7), s(8), s(9), s(10), &
         out(1), out(2), out(3), out(4), out(5))
  end subroutine ref_riemannsolver

  ! riemann (Newton) alone: io(1) = initial pstar guess -> pstar, io(2) = vstar.
  subroutine ref_riemann(io, pl, pr, rl, rr, vl, vr) bind(C, name="ref_riemann")
    real(c_double), intent(inout) :: io(2)
    real(c_double), value :: pl, pr, rl, rr, vl, vr
    real(c_double) :: a, b, c, d, e, f
    external :: riemann
    a = pl; b = pr; c = rl; d = rr; e = vl; f = vr
    call riemann(io(1), io(2), a, b, c, d, e, f)
  end subroutine ref_riemann

  ! integrate on one cell: f(10) = pm1,pm2,rhom1,rhom2,vxm1,vxm2,vym1,vym2,
  ! vzm1,vzm2; st(5) = pres,rho,vx,vy,vz updated in place; e(3) = etot,eint,ghalf.
  subroutine ref_integrate(f, st, e) bind(C, name="ref_integrate")
    real(c_double), intent(in) :: f(10), e(3)
    real(c_double), intent(inout) :: st(5)
    real(c_double) :: ff(10), ee(3)
    external :: integrate
    ff = f; ee = e
    call integrate(ff(1), ff(2), ff(3), ff(4), ff(5), ff(6), ff(7), ff(8), ff(9), ff(10), &
         st(1), st(2), st(3), st(4), st(5), ee(1), ee(2), ee(3))
  end subroutine ref_integrate

  ! The N-body set-up kick and drift read (hydropad/init_sys.F90:75-100 on one
  ! rank): the mesh sizes, one rank per dimension, boxsize (FP32 ratios, as
  ! there) and the periodic Cartesian communicator MPI_CART_RANK needs.
  subroutine ref_nbody_setup(ngx, ngy, ngz, nb) bind(C, name="ref_nbody_setup")
    use dimension
    use mpi_inc
    integer(c_int), value :: ngx, ngy, ngz, nb
    integer :: d(3), e
    logical :: per(3)
    call ref_mpi_init()
    ngridx = ngx; ngridy = ngy; ngridz = ngz
    ngridxpe = ngx; ngridype = ngy; ngridzpe = ngz
    npesx = 1; npesy = 1; npesz = 1
    nbound = nb
    if (.not. allocated(boxsize)) allocate(boxsize(3))
    boxsize(1) = real(ngridx)/real(max(ngridx,ngridy,ngridz))
    boxsize(2) = real(ngridy)/real(max(ngridx,ngridy,ngridz))
    boxsize(3) = real(ngridz)/real(max(ngridx,ngridy,ngridz))
    if (COMM_CART == 0) then
      d = 1; per = .true.
      call MPI_Cart_create(MPI_COMM_WORLD, 3, d, per, 0, COMM_CART, e)
    end if
  end subroutine ref_nbody_setup

  ! SUBROUTINE kick (nbody/kick.F90:1-133) on rank 0 with the module scalars
  ! dat and dtat; ppos/pvel are (3,np), gforce (3,nx,ny,nz).
  subroutine ref_kick(np, ppos, pvel, nx, ny, nz, gforce, datv, dtatv) bind(C, name="ref_kick")
    use scalar
    integer(c_int), value :: np, nx, ny, nz
    real(c_double), intent(in) :: ppos(3, np), gforce(3, nx, ny, nz)
    real(c_double), intent(inout) :: pvel(3, np)
    real(c_double), value :: datv, dtatv
    integer :: n, ix, iy, iz, keep
    external :: kick
    n = np; ix = nx; iy = ny; iz = nz
    dat = datv; dtat = dtatv
    keep = mype
    mype = 0
    call kick(n, ppos, pvel, ix, iy, iz, gforce)
    mype = keep
  end subroutine ref_kick

  ! SUBROUTINE drift (nbody/drift.F90:1-36) with the module scalar rdtath.
  subroutine ref_drift(np, ppos, pvel, rdtathv) bind(C, name="ref_drift")
    use scalar
    integer(c_int), value :: np
    real(c_double), intent(inout) :: ppos(3, np), pvel(3, np)
    real(c_double), value :: rdtathv
    integer :: n
    external :: drift
    n = np
    rdtath = rdtathv
    call drift(n, ppos, pvel)
  end subroutine ref_drift

end module hp_ref_harness
