! hydropad_driver.F90 -- a time loop shaped like hydropad/main.F90:100-121 and
! hydropad/evolve_sys.F90 (nrot schedule, ppm_lxlylz, speed, timedet) over
! the GPU hydro step, driven through the ISO_C_BINDING module.  Sets up the
! planar Sod shock tube (rho,p = 1,1 for cell centre x < 0.5, else 0.125,0.1),
! runs it, and writes the final fields in the reference's dump format
! (hydropad/main.F90:125-136 + io/write_var_mpi.F90: rho.bin, pressure.bin,
! vx.bin, vy.bin, vz.bin; REAL(4) interior, Fortran order, no header).
!
!   hydropad_driver N NSTEPS [fp64|mixed|exact] [OUTDIR] [resident|sweeps|cells|nbody] [NEWTON_LIMIT] [...]
!
! NEWTON_LIMIT (default 0 = the reference's 100) is the Riemann Newton
! iteration count past which the run stops as hydro/riemann.F90:60-65 does.
! Mode "cells" runs no time loop: it calls the integrate_gpu drop-in on one
! cold cell with equal face fluxes for the internal energies given as the
! remaining arguments and prints each new pressure.
! Mode "nbody" runs the N-body drop-ins over NSTEPS steps of kick, drift,
! kick (hydropad/evolve_sys.F90:63-84) with dtat = 0.015, dat = 0.2,
! rdtath = 0.04, on gforce.bin (3,nx,ny,nz) and NP particles ppos.bin,
! pvel.bin (3,NP) REAL(8) read from OUTDIR (NP is argument 7): once with the
! per-call wrappers kick_gpu / drift_gpu (host arrays in and out, as the
! reference calls kick and drift) into ppos_call.bin / pvel_call.bin, once
! device-resident (particles attached, nbody_step_gpu per step) into
! ppos_res.bin / pvel_res.bin, and prints the particles-outside counter.
! "resident" keeps the state on the GPU for the whole run (attach /
! ppm_lxlylz_gpu / speed_gpu / detach); "sweeps" calls the per-sweep drop-in
! ppmsolver_gpu with host arrays exactly where hydro/ppm_lxlylz.F90 calls
! ppmsolver, and computes speed on the host (hydropad/speed.F90:22-31).
program hydropad_driver
  use iso_c_binding
  use hydropad_gpu
  implicit none
  integer :: n, nsteps, nb, nx, ny, nz, nstep, nrot, i, isweep
  integer, parameter :: order(3, 6) = reshape([1,2,3, 3,2,1, 2,3,1, 1,3,2, 3,1,2, 2,1,3], [3, 6])  ! ppm_lxlylz.F90
  real(c_double), allocatable :: p3d(:,:,:), rho3d(:,:,:), vx3d(:,:,:), vy3d(:,:,:), vz3d(:,:,:)
  real(c_double), allocatable :: gforce(:,:,:,:)
  real(c_double) :: cho3d(1), nes3d(1)
  character(len=16) :: mode
  real(c_double) :: gamma, gamma1, rgamma, rgamma1, m, rm, gf, dx, rdx, dt, t, cour, velmax, x, tstop
  real(c_double) :: flatvalue, mingradflat, eta1, eta2, dmax
  character(len=256) :: arg, outdir

  n = 64; nsteps = 10; outdir = '.'; mode = 'resident'
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg); read(arg, *) n
  end if
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg); read(arg, *) nsteps
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    if (trim(arg) == 'mixed') call hydropad_gpu_set_precision(HP_MIXED)
    if (trim(arg) == 'exact') call hydropad_gpu_set_precision(HP_FP64_EXACT)
  end if
  if (command_argument_count() >= 4) call get_command_argument(4, outdir)
  if (command_argument_count() >= 5) call get_command_argument(5, mode)
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg, *) i
    call hydropad_gpu_set_newton_limit(i)
  end if

  nb = 3
  nx = n + 2*nb; ny = nx; nz = nx
  allocate(p3d(nx,ny,nz), rho3d(nx,ny,nz), vx3d(nx,ny,nz), vy3d(nx,ny,nz), vz3d(nx,ny,nz))
  vx3d = 0.0d0; vy3d = 0.0d0; vz3d = 0.0d0
  do i = 1, nx
    x = (dble(i - nb) - 0.5d0) / dble(n)
    if (x < 0.5d0) then
      rho3d(i,:,:) = 1.0d0;   p3d(i,:,:) = 1.0d0
    else
      rho3d(i,:,:) = 0.125d0; p3d(i,:,:) = 0.1d0
    end if
  end do
  call fill_periodic(p3d); call fill_periodic(rho3d)

  ! derived constants, hydropad/initialize_parameters.F90:23-24,64-69
  gamma = 5.0d0/3.0d0
  rgamma = 1.0d0/gamma
  gamma1 = (gamma+1.0d0)/(gamma-1.0d0)
  rgamma1 = 1.0d0/(gamma-1)
  m = (gamma-1.0d0)/(2.0d0*gamma)
  rm = 1.0d0/m
  gf = (gamma+1.0d0)/(2.0d0*gamma)
  dx = 1.0/real(n)
  rdx = 1.0d0/dx
  flatvalue = 0.15d0; mingradflat = 0.25d0; eta1 = 1.0d-3; eta2 = 0.1d0; dmax = 2.0d10
  cour = 0.8d0; tstop = 1.0d30; t = 0.0d0; dt = 0.0d0

  call hydropad_gpu_set_scalars(mingradflat, flatvalue, gamma, 1.0d0, m, rgamma1, rm, rgamma, dmax, &
                                rdx, dt, dt, dx, 0.0d0, 1.0d0, gamma1, gf, eta1, eta2)
  if (trim(mode) == 'cells') then
    call integrate_cells()
    call hydropad_gpu_finalize()
    stop
  end if
  if (trim(mode) == 'nbody') then
    call nbody_run()
    call hydropad_gpu_finalize()
    stop
  end if
  if (trim(mode) == 'sweeps') then
    allocate(gforce(3, nx, ny, nz))
    gforce = 0.0d0
    call timedet(host_speed())          ! the first dt from the initial state's CFL condition
    do nstep = 1, nsteps                ! hydropad/main.F90:100-121
      nrot = mod(nstep, 6) + 1          ! hydropad/evolve_sys.F90:30-31
      if (nstep == 1) nrot = 1
      call hydropad_gpu_set_scalars(mingradflat, flatvalue, gamma, 1.0d0, m, rgamma1, rm, rgamma, dmax, &
                                    rdx, dt / 1.0d0, dt, dx, 0.0d0, 1.0d0, gamma1, gf, eta1, eta2)
      do isweep = 1, 3                  ! hydro/ppm_lxlylz.F90: ppmsolver + exchange_mesh per sweep
        call ppmsolver_gpu(p3d, rho3d, vx3d, vy3d, vz3d, cho3d, nes3d, gforce, nx, ny, nz, nb, &
                           order(isweep, nrot))
      end do
      call timedet(host_speed())
    end do
  else
    call hydropad_gpu_attach(p3d, rho3d, vx3d, vy3d, vz3d, nx, ny, nz, nb)
    call speed_gpu(velmax)
    call timedet(velmax)                ! the first dt from the initial state's CFL condition
    do nstep = 1, nsteps                ! hydropad/main.F90:100-121
      nrot = mod(nstep, 6) + 1          ! hydropad/evolve_sys.F90:30-31
      if (nstep == 1) nrot = 1
      call hydropad_gpu_set_scalars(mingradflat, flatvalue, gamma, 1.0d0, m, rgamma1, rm, rgamma, dmax, &
                                    rdx, dt / 1.0d0, dt, dx, 0.0d0, 1.0d0, gamma1, gf, eta1, eta2)
      call ppm_lxlylz_gpu(nrot)
      call speed_gpu(velmax)
      call timedet(velmax)
    end do
    call hydropad_gpu_detach(p3d, rho3d, vx3d, vy3d, vz3d)
  end if
  call write_var(rho3d, 'rho.bin'); call write_var(p3d, 'pressure.bin')
  call write_var(vx3d, 'vx.bin'); call write_var(vy3d, 'vy.bin'); call write_var(vz3d, 'vz.bin')
  write(*, '(a,i0,a,i0,a,es24.16,a,es24.16,a,es24.16)') 'hydropad_driver: N=', n, ' steps=', nsteps, &
      ' t=', t, ' dt=', dt, ' sum(rho)=', sum(rho3d(nb+1:nx-nb, nb+1:ny-nb, nb+1:nz-nb))
  call hydropad_gpu_finalize()

contains

  ! integrate_gpu (SUBROUTINE integrate, hydro/integrate.F90:1-2) on a cold
  ! cell, e/E ~ 1e-4 < eta1, with the caller's internal energy eint taken from
  ! the command line (arguments 7, 8, ...).
  subroutine integrate_cells()
    integer :: a
    real(c_double) :: eint, etot, pres, rho, vx, vy, vz
    do a = 7, command_argument_count()
      call get_command_argument(a, arg); read(arg, *) eint
      pres = 1.0d-4; rho = 1.0d0; vx = 2.0d0; vy = 0.0d0; vz = 0.0d0
      etot = pres * 1.5d0 + 0.5d0 * rho * vx * vx
      call integrate_gpu(0.3d0, 0.3d0, 1.1d0, 1.1d0, 2.0d0, 2.0d0, 0.1d0, 0.1d0, -0.2d0, -0.2d0, &
                         pres, rho, vx, vy, vz, etot, eint, 0.0d0)
      write(*, '(a,es26.17e3,a,es26.17e3)') 'integrate: eint=', eint, ' pres=', pres
    end do
  end subroutine integrate_cells

  ! kick, drift, kick per step through the per-call and the resident drop-ins
  subroutine nbody_run()
    integer :: np, s
    real(c_double), allocatable :: g(:,:,:,:), x0(:,:), v0(:,:), x(:,:), v(:,:)
    type(hp_counters_t) :: cnt
    call get_command_argument(7, arg); read(arg, *) np
    allocate(g(3, nx, ny, nz), x0(3, np), v0(3, np), x(3, np), v(3, np))
    call read_raw(g, size(g), 'gforce.bin')
    call read_raw(x0, size(x0), 'ppos.bin')
    call read_raw(v0, size(v0), 'pvel.bin')
    call hydropad_gpu_set_scalars(mingradflat, flatvalue, gamma, 1.0d0, m, rgamma1, rm, rgamma, dmax, &
                                  rdx, 0.04d0, 0.0d0, dx, 0.2d0, 1.0d0, gamma1, gf, eta1, eta2)
    call hydropad_gpu_set_dtat(0.015d0)
    x = x0; v = v0
    do s = 1, nsteps                      ! hydropad/evolve_sys.F90:63-84, one rank
      call kick_gpu(np, x, v, nx, ny, nz, g)
      call drift_gpu(np, x, v)
      call kick_gpu(np, x, v, nx, ny, nz, g)
    end do
    call write_raw(x, size(x), 'ppos_call.bin'); call write_raw(v, size(v), 'pvel_call.bin')
    call hydropad_gpu_attach(p3d, rho3d, vx3d, vy3d, vz3d, nx, ny, nz, nb)
    call hydropad_gpu_set_gforce(g)
    call hydropad_gpu_particles_attach(np, x0, v0)
    do s = 1, nsteps
      call nbody_step_gpu()
    end do
    call hydropad_gpu_particles_detach(x, v)
    call write_raw(x, size(x), 'ppos_res.bin'); call write_raw(v, size(v), 'pvel_res.bin')
    call hydropad_gpu_counters(cnt)
    write(*, '(a,i0)') 'hydropad_driver: particles_outside=', cnt%particles_outside
  end subroutine nbody_run

  subroutine read_raw(a, n, name)
    integer, intent(in) :: n
    real(c_double), intent(out) :: a(n)
    character(len=*), intent(in) :: name
    integer :: u
    open(newunit=u, file=trim(outdir)//'/'//name, access='stream', form='unformatted', status='old')
    read(u) a
    close(u)
  end subroutine read_raw

  subroutine write_raw(a, n, name)
    integer, intent(in) :: n
    real(c_double), intent(in) :: a(n)
    character(len=*), intent(in) :: name
    integer :: u
    open(newunit=u, file=trim(outdir)//'/'//name, access='stream', form='unformatted', status='replace')
    write(u) a
    close(u)
  end subroutine write_raw

  ! CFL maximum on the host, hydropad/speed.F90:22-31
  real(c_double) function host_speed()
    integer :: ii, jj, kk
    real(c_double) :: cs
    host_speed = 0.0d0
    do kk = nb + 1, nz - nb
      do jj = nb + 1, ny - nb
        do ii = nb + 1, nx - nb
          cs = sqrt(gamma * p3d(ii, jj, kk) / rho3d(ii, jj, kk))
          host_speed = max(host_speed, max(abs(vx3d(ii, jj, kk)), abs(vy3d(ii, jj, kk)), abs(vz3d(ii, jj, kk))) + cs)
        end do
      end do
    end do
  end function host_speed

  ! hydropad/functions.F90:2-35 (non-cosmological)
  subroutine timedet(vmax)
    real(c_double), intent(in) :: vmax
    real(c_double) :: told, rdtc
    told = t
    rdtc = vmax / (cour * dx)
    if (vmax == 0.0d0) then
      dt = 0.1d0
    else
      dt = 1.0d0 / rdtc
    end if
    t = t + dt
    if (t > tstop) then
      t = tstop
      dt = t - told
    end if
  end subroutine timedet

  ! periodic ghosts, z then x then y (hydropad/exchange_mesh.F90:47-123)
  subroutine fill_periodic(a)
    real(c_double), intent(inout) :: a(nx, ny, nz)
    integer :: k
    do k = 1, nb
      a(:, :, k) = a(:, :, nz - 2*nb + k); a(:, :, nz - nb + k) = a(:, :, k + nb)
    end do
    do k = 1, nb
      a(k, :, :) = a(nx - 2*nb + k, :, :); a(nx - nb + k, :, :) = a(k + nb, :, :)
    end do
    do k = 1, nb
      a(:, k, :) = a(:, ny - 2*nb + k, :); a(:, ny - nb + k, :) = a(:, k + nb, :)
    end do
  end subroutine fill_periodic

  ! REAL(4) interior, Fortran order, no header (io/write_var_mpi.F90:55-65 on one rank)
  subroutine write_var(a, name)
    real(c_double), intent(in) :: a(nx, ny, nz)
    character(len=*), intent(in) :: name
    integer :: u
    open(newunit=u, file=trim(outdir)//'/'//name, access='stream', form='unformatted', status='replace')
    write(u) real(a(nb+1:nx-nb, nb+1:ny-nb, nb+1:nz-nb))
    close(u)
  end subroutine write_var

end program hydropad_driver
