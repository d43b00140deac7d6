! hydropad_gpu.F90 -- ISO_C_BINDING interface of include/hydropad_gpu.h for
! Fortran drivers, plus drop-in wrappers that keep the dummy-argument lists of
! the reference's hydro subroutines:
!
!   ppmsolver_gpu(p3d,rho3d,vx3d,vy3d,vz3d,cho3d,nes3d,gforce,nx,ny,nz,nbound,ndirection)
!       = SUBROUTINE ppmsolver, STENCIL argument list (hydro/ppmsolver.F90:1-5)
!   integrate_gpu(pm1,pm2,rhom1,rhom2,vxm1,vxm2,vym1,vym2,vzm1,vzm2,
!                 pres,rho,vx,vy,vz,etot,eint,ghalf)
!       = SUBROUTINE integrate (hydro/integrate.F90:1-2)
!   dualenergy_gpu(etotnew,rhonew,vx,vy,vz,eintnew,eint,pres,pnew)
!       = the dual-energy selection inlined at hydro/integrate.F90:196-216
!   hydropad_gpu_set_scalars(mingradflat,flatvalue,gamma,rat,m,rgamma1,rm,rgamma,
!                 dmax,rdx,rdtath,dt,dx,dat,at,gamma1,gf,eta1,eta2)
!       = the scalar push before every sweep (hydro/ppmsolver.F90:68-69), same order
!
!   kick_gpu(np,ppos,pvel,ngx,ngy,ngz,gforce)
!       = SUBROUTINE kick (nbody/kick.F90:1)
!   drift_gpu(np,ppos,pvel)
!       = SUBROUTINE drift (nbody/drift.F90:1)
!   hydropad_gpu_set_dtat(dtat)
!       = the module scalar dtat the kicks read (hydropad/scalar.F90:18)
!
! and a device-resident form for an evolve_sys-shaped time loop:
!   hydropad_gpu_attach(p3d,rho3d,vx3d,vy3d,vz3d,nx,ny,nz,nbound)   upload once
!   hydropad_gpu_set_gforce(gforce)     gravity array, once (initialize_gravity)
!   hydropad_gpu_particles_attach(np,ppos,pvel)                     upload once
!   ppm_lxlylz_gpu(nrot)                (hydro/ppm_lxlylz.F90, 3 sweeps + ghost fills)
!   nbody_step_gpu()                    (kick, drift, kick: evolve_sys.F90:63-84)
!   speed_gpu(velmax)                   (hydropad/speed.F90, particles included)
!   hydropad_gpu_particles_detach(ppos,pvel)                        download once
!   hydropad_gpu_detach(p3d,rho3d,vx3d,vy3d,vz3d)                  download once
module hydropad_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: HP_FP64 = 0, HP_MIXED = 1, HP_FP64_EXACT = 2

  type, bind(C), public :: hp_params_t
    real(c_double) :: gamma, gamma1, rgamma, rgamma1, m, rm, gf
    real(c_double) :: dt, dx, rdx, rdtath, at, dat, rat
    real(c_double) :: flatvalue, mingradflat, eta1, eta2, dmax
  end type hp_params_t

  type, bind(C), public :: hp_counters_t
    integer(c_int64_t) :: newton_iters_max, newton_nonconv, neg_state_clamps
    integer(c_int64_t) :: particles_outside
    integer(c_int64_t) :: reserved(4)
  end type hp_counters_t

  public :: hp_params_init, hp_create, hp_destroy, hp_set_precision, hp_upload, hp_download, &
            hp_set_params, hp_ppmsolver, hp_exchange_mesh, hp_ppm_lxlylz, hp_speed, hp_evolve, &
            hp_status, hp_set_newton_limit, hp_integrate_cell, hp_dualenergy_cell
  public :: hydropad_gpu_set_scalars, hydropad_gpu_set_precision, hydropad_gpu_set_newton_limit, &
            ppmsolver_gpu, integrate_gpu, &
            dualenergy_gpu, hydropad_gpu_attach, ppm_lxlylz_gpu, speed_gpu, exchange_mesh_gpu, &
            hydropad_gpu_detach, hydropad_gpu_finalize, hydropad_gpu_check, &
            hydropad_gpu_set_dtat, kick_gpu, drift_gpu, hydropad_gpu_set_gforce, &
            hydropad_gpu_particles_attach, nbody_step_gpu, hydropad_gpu_particles_detach, &
            hydropad_gpu_counters

  interface
    integer(c_int) function hp_params_init(gamma, ngrid_max, dt, out) bind(C, name="hp_params_init")
      import :: c_int, c_double, hp_params_t
      real(c_double), value :: gamma, dt
      integer(c_int), value :: ngrid_max
      type(hp_params_t), intent(out) :: out
    end function
    integer(c_int) function hp_create(nx, ny, nz, nbound, precision, ctx) bind(C, name="hp_create")
      import :: c_int, c_ptr
      integer(c_int), value :: nx, ny, nz, nbound, precision
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine hp_destroy(ctx) bind(C, name="hp_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    integer(c_int) function hp_set_precision(ctx, precision) bind(C, name="hp_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
    end function
    integer(c_int) function hp_upload(ctx, p, rho, vx, vy, vz, gforce) bind(C, name="hp_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: p(*), rho(*), vx(*), vy(*), vz(*)
      type(c_ptr), value :: gforce
    end function
    integer(c_int) function hp_download(ctx, p, rho, vx, vy, vz) bind(C, name="hp_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: p(*), rho(*), vx(*), vy(*), vz(*)
    end function
    integer(c_int) function hp_set_params(ctx, prm) bind(C, name="hp_set_params")
      import :: c_int, c_ptr, hp_params_t
      type(c_ptr), value :: ctx
      type(hp_params_t), intent(in) :: prm
    end function
    integer(c_int) function hp_ppmsolver(ctx, ndirection) bind(C, name="hp_ppmsolver")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ndirection
    end function
    integer(c_int) function hp_exchange_mesh(ctx, direction) bind(C, name="hp_exchange_mesh")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: direction
    end function
    integer(c_int) function hp_ppm_lxlylz(ctx, nrot) bind(C, name="hp_ppm_lxlylz")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrot
    end function
    integer(c_int) function hp_speed(ctx, velmax) bind(C, name="hp_speed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: velmax
    end function
    integer(c_int) function hp_evolve(ctx, first_step, nsteps, cour, tstop, t, dt) bind(C, name="hp_evolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: first_step, nsteps
      real(c_double), value :: cour, tstop
      real(c_double), intent(inout) :: t, dt
    end function
    integer(c_int) function hp_upload_gforce(ctx, gforce) bind(C, name="hp_upload_gforce")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, gforce
    end function
    integer(c_int) function hp_particles_upload(ctx, np, ppos, pvel, dtat) bind(C, name="hp_particles_upload")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: np
      real(c_double), intent(in) :: ppos(*), pvel(*)
      real(c_double), value :: dtat
    end function
    integer(c_int) function hp_particles_download(ctx, ppos, pvel) bind(C, name="hp_particles_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ppos(*), pvel(*)
    end function
    integer(c_int) function hp_kick(ctx, dtat) bind(C, name="hp_kick")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dtat
    end function
    integer(c_int) function hp_drift(ctx) bind(C, name="hp_drift")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hp_set_newton_limit(ctx, nmax) bind(C, name="hp_set_newton_limit")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nmax
    end function
    integer(c_int) function hp_status(ctx, c) bind(C, name="hp_status")
      import :: c_int, c_ptr, hp_counters_t
      type(c_ptr), value :: ctx
      type(hp_counters_t), intent(out) :: c
    end function
    integer(c_int) function hp_integrate_cell(prm, pm1, pm2, rhom1, rhom2, vxm1, vxm2, vym1, vym2, &
                                              vzm1, vzm2, pres, rho, vx, vy, vz, etot, eint, ghalf) &
                                              bind(C, name="hp_integrate_cell")
      import :: c_int, c_double, hp_params_t
      type(hp_params_t), intent(in) :: prm
      real(c_double), intent(in) :: pm1, pm2, rhom1, rhom2, vxm1, vxm2, vym1, vym2, vzm1, vzm2
      real(c_double), intent(inout) :: pres, rho, vx, vy, vz
      real(c_double), intent(in) :: etot, eint, ghalf
    end function
    integer(c_int) function hp_dualenergy_cell(prm, etotnew, rhonew, vx, vy, vz, eintnew_in, eint_old, &
                                               pres_old, pnew, eintnew) bind(C, name="hp_dualenergy_cell")
      import :: c_int, c_double, hp_params_t
      type(hp_params_t), intent(in) :: prm
      real(c_double), value :: etotnew, rhonew, vx, vy, vz, eintnew_in, eint_old, pres_old
      real(c_double), intent(out) :: pnew, eintnew
    end function
  end interface

  ! state of the drop-in wrappers: one context, the scalars of the last push
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: cnx = 0, cny = 0, cnz = 0, cnb = 0
  integer(c_int), save :: cprec = HP_FP64
  type(hp_params_t), save :: prm
  integer(c_int), save :: cnewt = 0
  real(c_double), save :: cdtat = 0.0d0

contains

  ! Stop with a message on a nonzero C-ABI return code.  HP_E_NEWTON (-4) is
  ! the reference's own stop on a Riemann Newton iteration that does not
  ! converge, with its message (hydro/riemann.F90:60-65).
  subroutine hydropad_gpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc == -4_c_int) then
      write(*,*) 'convergenza non raggiunta'
      stop
    end if
    if (rc /= 0) then
      write(*,*) 'hydropad_gpu: ', what, ' failed, code ', rc
      stop 1
    end if
  end subroutine hydropad_gpu_check

  ! The 19 scalars in the order the reference pushes them (ppmsolver.F90:68-69).
  subroutine hydropad_gpu_set_scalars(mingradflat, flatvalue, gamma, rat, m, rgamma1, rm, rgamma, dmax, &
                                      rdx, rdtath, dt, dx, dat, at, gamma1, gf, eta1, eta2)
    real(c_double), intent(in) :: mingradflat, flatvalue, gamma, rat, m, rgamma1, rm, rgamma, dmax
    real(c_double), intent(in) :: rdx, rdtath, dt, dx, dat, at, gamma1, gf, eta1, eta2
    prm%mingradflat = mingradflat; prm%flatvalue = flatvalue; prm%gamma = gamma; prm%rat = rat
    prm%m = m; prm%rgamma1 = rgamma1; prm%rm = rm; prm%rgamma = rgamma; prm%dmax = dmax
    prm%rdx = rdx; prm%rdtath = rdtath; prm%dt = dt; prm%dx = dx; prm%dat = dat; prm%at = at
    prm%gamma1 = gamma1; prm%gf = gf; prm%eta1 = eta1; prm%eta2 = eta2
    if (c_associated(ctx)) call hydropad_gpu_check(hp_set_params(ctx, prm), 'hp_set_params')
  end subroutine hydropad_gpu_set_scalars

  subroutine hydropad_gpu_set_precision(precision)
    integer(c_int), intent(in) :: precision
    cprec = precision
    if (c_associated(ctx)) call hydropad_gpu_check(hp_set_precision(ctx, cprec), 'hp_set_precision')
  end subroutine hydropad_gpu_set_precision

  ! Newton iterations after which the run stops (the reference's nnewtmax = 100,
  ! hydro/riemann.F90:51; 0 restores it).
  subroutine hydropad_gpu_set_newton_limit(nmax)
    integer, intent(in) :: nmax
    cnewt = int(nmax, c_int)
    if (c_associated(ctx)) call hydropad_gpu_check(hp_set_newton_limit(ctx, cnewt), 'hp_set_newton_limit')
  end subroutine hydropad_gpu_set_newton_limit

  subroutine ensure_ctx(nx, ny, nz, nbound)
    integer, intent(in) :: nx, ny, nz, nbound
    if (c_associated(ctx) .and. (nx /= cnx .or. ny /= cny .or. nz /= cnz .or. nbound /= cnb)) then
      call hp_destroy(ctx)
      ctx = c_null_ptr
    end if
    if (.not. c_associated(ctx)) then
      call hydropad_gpu_check(hp_create(int(nx, c_int), int(ny, c_int), int(nz, c_int), &
                                        int(nbound, c_int), cprec, ctx), 'hp_create')
      cnx = nx; cny = ny; cnz = nz; cnb = nbound
      call hydropad_gpu_check(hp_set_newton_limit(ctx, cnewt), 'hp_set_newton_limit')
    end if
    call hydropad_gpu_check(hp_set_params(ctx, prm), 'hp_set_params')
  end subroutine ensure_ctx

  ! Drop-in for SUBROUTINE ppmsolver (STENCIL arguments, hydro/ppmsolver.F90:1-5):
  ! host arrays in, one sweep on the GPU, host arrays out (like the
  ! reference's acc update device/host around each sweep).  cho3d and nes3d
  ! are unused, as under STENCIL; an all-zero gforce selects the no-gravity kernels.
  subroutine ppmsolver_gpu(p3d, rho3d, vx3d, vy3d, vz3d, cho3d, nes3d, gforce, nx, ny, nz, nbound, ndirection)
    integer, intent(in) :: nx, ny, nz, nbound, ndirection
    real(c_double), intent(inout) :: p3d(nx, ny, nz), rho3d(nx, ny, nz), vx3d(nx, ny, nz)
    real(c_double), intent(inout) :: vy3d(nx, ny, nz), vz3d(nx, ny, nz)
    real(c_double), intent(in) :: cho3d(*), nes3d(*)
    real(c_double), intent(in), target :: gforce(3, nx, ny, nz)
    type(c_ptr) :: g
    call ensure_ctx(nx, ny, nz, nbound)
    g = c_null_ptr
    if (any(gforce /= 0.0d0)) g = c_loc(gforce)
    call hydropad_gpu_check(hp_upload(ctx, p3d, rho3d, vx3d, vy3d, vz3d, g), 'hp_upload')
    call hydropad_gpu_check(hp_ppmsolver(ctx, int(ndirection, c_int)), 'hp_ppmsolver')
    call hydropad_gpu_check(hp_download(ctx, p3d, rho3d, vx3d, vy3d, vz3d), 'hp_download')
  end subroutine ppmsolver_gpu

  ! Drop-in for SUBROUTINE integrate (hydro/integrate.F90), run on the device.
  subroutine integrate_gpu(pm1, pm2, rhom1, rhom2, vxm1, vxm2, vym1, vym2, vzm1, vzm2, &
                           pres, rho, vx, vy, vz, etot, eint, ghalf)
    real(c_double), intent(in) :: pm1, pm2, rhom1, rhom2, vxm1, vxm2, vym1, vym2, vzm1, vzm2
    real(c_double), intent(inout) :: pres, rho, vx, vy, vz
    real(c_double), intent(in) :: etot, eint, ghalf
    call hydropad_gpu_check(hp_integrate_cell(prm, pm1, pm2, rhom1, rhom2, vxm1, vxm2, vym1, vym2, &
                                              vzm1, vzm2, pres, rho, vx, vy, vz, etot, eint, ghalf), &
                            'hp_integrate_cell')
  end subroutine integrate_gpu

  ! The dual-energy selection of hydro/integrate.F90:196-216, run on the device.
  subroutine dualenergy_gpu(etotnew, rhonew, vx, vy, vz, eintnew, eint, pres, pnew)
    real(c_double), intent(in) :: etotnew, rhonew, vx, vy, vz, eint, pres
    real(c_double), intent(inout) :: eintnew
    real(c_double), intent(out) :: pnew
    real(c_double) :: ein
    ein = eintnew
    call hydropad_gpu_check(hp_dualenergy_cell(prm, etotnew, rhonew, vx, vy, vz, ein, eint, pres, &
                                               pnew, eintnew), 'hp_dualenergy_cell')
  end subroutine dualenergy_gpu

  ! Device-resident time loop: upload the fields once ...
  subroutine hydropad_gpu_attach(p3d, rho3d, vx3d, vy3d, vz3d, nx, ny, nz, nbound)
    integer, intent(in) :: nx, ny, nz, nbound
    real(c_double), intent(in) :: p3d(nx, ny, nz), rho3d(nx, ny, nz), vx3d(nx, ny, nz)
    real(c_double), intent(in) :: vy3d(nx, ny, nz), vz3d(nx, ny, nz)
    call ensure_ctx(nx, ny, nz, nbound)
    call hydropad_gpu_check(hp_upload(ctx, p3d, rho3d, vx3d, vy3d, vz3d, c_null_ptr), 'hp_upload')
  end subroutine hydropad_gpu_attach

  ! ... one full step in the order nrot (hydro/ppm_lxlylz.F90:9-221) ...
  subroutine ppm_lxlylz_gpu(nrot)
    integer, intent(in) :: nrot
    call hydropad_gpu_check(hp_set_params(ctx, prm), 'hp_set_params')
    call hydropad_gpu_check(hp_ppm_lxlylz(ctx, int(nrot, c_int)), 'hp_ppm_lxlylz')
  end subroutine ppm_lxlylz_gpu

  ! ... the CFL maximum of hydropad/speed.F90 ...
  subroutine speed_gpu(velmax)
    real(c_double), intent(out) :: velmax
    call hydropad_gpu_check(hp_speed(ctx, velmax), 'hp_speed')
  end subroutine speed_gpu

  subroutine exchange_mesh_gpu()
    call hydropad_gpu_check(hp_exchange_mesh(ctx, 0_c_int), 'hp_exchange_mesh')
  end subroutine exchange_mesh_gpu

  ! ... and download at the end.
  subroutine hydropad_gpu_detach(p3d, rho3d, vx3d, vy3d, vz3d)
    real(c_double), intent(out) :: p3d(*), rho3d(*), vx3d(*), vy3d(*), vz3d(*)
    call hydropad_gpu_check(hp_download(ctx, p3d, rho3d, vx3d, vy3d, vz3d), 'hp_download')
  end subroutine hydropad_gpu_detach

  ! The kicks' time step, the reference's module scalar dtat
  ! (hydropad/scalar.F90:18, set at initialize_parameters.F90:174,193).
  subroutine hydropad_gpu_set_dtat(dtat)
    real(c_double), intent(in) :: dtat
    cdtat = dtat
  end subroutine hydropad_gpu_set_dtat

  ! Drop-in for SUBROUTINE kick (nbody/kick.F90:1, one rank): host particles and
  ! gravity array in, v <- v (1 - dat dtat/2) + F dtat/2 on the GPU, velocities
  ! out.  ngx, ngy, ngz are the array dimensions (ghosts included); the grid's
  ! nbound is the attached context's, else the reference's 3.
  subroutine kick_gpu(np, ppos, pvel, ngx, ngy, ngz, gforce)
    integer, intent(in) :: np, ngx, ngy, ngz
    real(c_double), intent(in) :: ppos(3, np)
    real(c_double), intent(inout) :: pvel(3, np)
    real(c_double), intent(in), target :: gforce(3, ngx, ngy, ngz)
    real(c_double), allocatable :: pos(:, :)
    integer :: nb
    nb = 3
    if (c_associated(ctx) .and. ngx == cnx .and. ngy == cny .and. ngz == cnz) nb = cnb
    call ensure_ctx(ngx, ngy, ngz, nb)
    call hydropad_gpu_check(hp_upload_gforce(ctx, c_loc(gforce)), 'hp_upload_gforce')
    call hydropad_gpu_check(hp_particles_upload(ctx, int(np, c_int64_t), ppos, pvel, cdtat), 'hp_particles_upload')
    call hydropad_gpu_check(hp_kick(ctx, cdtat), 'hp_kick')
    allocate(pos(3, np))
    call hydropad_gpu_check(hp_particles_download(ctx, pos, pvel), 'hp_particles_download')
  end subroutine kick_gpu

  ! Drop-in for SUBROUTINE drift (nbody/drift.F90:1, one rank): x <- x + rdtath v
  ! wrapped once into the box, on the grid of the context a hydro call or
  ! kick_gpu set up (the box is [0, ngrid/max(ngrid)) per axis).
  subroutine drift_gpu(np, ppos, pvel)
    integer, intent(in) :: np
    real(c_double), intent(inout) :: ppos(3, np)
    real(c_double), intent(in) :: pvel(3, np)
    real(c_double), allocatable :: vel(:, :)
    if (.not. c_associated(ctx)) then
      write(*,*) 'hydropad_gpu: drift_gpu needs the grid of a context (a hydro call or kick_gpu first)'
      stop 1
    end if
    call hydropad_gpu_check(hp_set_params(ctx, prm), 'hp_set_params')
    call hydropad_gpu_check(hp_particles_upload(ctx, int(np, c_int64_t), ppos, pvel, cdtat), 'hp_particles_upload')
    call hydropad_gpu_check(hp_drift(ctx), 'hp_drift')
    allocate(vel(3, np))
    call hydropad_gpu_check(hp_particles_download(ctx, ppos, vel), 'hp_particles_download')
  end subroutine drift_gpu

  ! Device-resident N-body: the gravity array once ...
  subroutine hydropad_gpu_set_gforce(gforce)
    real(c_double), intent(in), target :: gforce(3, cnx, cny, cnz)
    call hydropad_gpu_check(hp_upload_gforce(ctx, c_loc(gforce)), 'hp_upload_gforce')
  end subroutine hydropad_gpu_set_gforce

  ! ... the particles once (the attached context's grid) ...
  subroutine hydropad_gpu_particles_attach(np, ppos, pvel)
    integer, intent(in) :: np
    real(c_double), intent(in) :: ppos(3, np), pvel(3, np)
    call hydropad_gpu_check(hp_particles_upload(ctx, int(np, c_int64_t), ppos, pvel, cdtat), 'hp_particles_upload')
  end subroutine hydropad_gpu_particles_attach

  ! ... kick, drift, kick of one step (hydropad/evolve_sys.F90:63-84) ...
  subroutine nbody_step_gpu()
    call hydropad_gpu_check(hp_set_params(ctx, prm), 'hp_set_params')
    call hydropad_gpu_check(hp_kick(ctx, cdtat), 'hp_kick')
    call hydropad_gpu_check(hp_drift(ctx), 'hp_drift')
    call hydropad_gpu_check(hp_kick(ctx, cdtat), 'hp_kick')
  end subroutine nbody_step_gpu

  ! ... and the particles back at the end.
  subroutine hydropad_gpu_particles_detach(ppos, pvel)
    real(c_double), intent(out) :: ppos(*), pvel(*)
    call hydropad_gpu_check(hp_particles_download(ctx, ppos, pvel), 'hp_particles_download')
  end subroutine hydropad_gpu_particles_detach

  ! The context's counters (Newton, clamps, particles outside the array).
  subroutine hydropad_gpu_counters(cnt)
    type(hp_counters_t), intent(out) :: cnt
    call hydropad_gpu_check(hp_status(ctx, cnt), 'hp_status')
  end subroutine hydropad_gpu_counters

  subroutine hydropad_gpu_finalize()
    if (c_associated(ctx)) call hp_destroy(ctx)
    ctx = c_null_ptr
  end subroutine hydropad_gpu_finalize

end module hydropad_gpu
