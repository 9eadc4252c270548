! picard_setup.f90 -- run configuration, derived constants and host particle
! sources for the Fortran host.
!
! The configuration is a namelist (&picard_run) instead of the reference's
! compile-time sections 1.1 / 2.1.0 (picard.f90:42-302); derive() evaluates the
! constants of picard.f90:333-466 for one domain; the sources follow
! ParticlesInitialize.f90 and BCinflow.f90 with random_number seeded
! explicitly (the reference's seeding, picard.f90:136-138, puts an
! uninitialised seed).
module picard_setup
  use iso_c_binding
  use picard_hip_iface
  implicit none

  real*8, parameter :: TAU = 6.2831853071795864769252867665590d0
  real*8, parameter :: EC = 1.60217646d-19, MP = 1.6726216d-27, ME = 9.1093819d-31
  real*8, parameter :: EPS0 = 8.854187817d-12, KB = 1.380650d-23
  real*8, parameter :: MU0 = 1.25663706143591729538505735331180d-6

  type run_config
    integer :: nx = 32, ny = 32, nz = 32, species = 2, steps = 10, maxit = 2**20, seed = 20261203
    logical :: xflow = .true., yflow = .false., zflow = .false.
    real*8 :: xmin = -160d0, xmax = 160d0, ymin = -160d0, ymax = 160d0, zmin = -160d0, zmax = 160d0
    real*8 :: dt_wpe = 0.1d0, maxerr = 1d-6
    real*8 :: mass(8) = [ME, MP, 18*MP, 4*MP, 44*MP, MP, MP, MP]
    real*8 :: density(8) = [0d0, 5d6, 0d0, 0d0, 0d0, 0d0, 0d0, 0d0]
    integer :: ppc(8) = [0, 8, 0, 0, 0, 0, 0, 0]
    real*8 :: kelvin(8) = [116045d0, 116045d0, 116045d0, 0d0, 0d0, 0d0, 0d0, 0d0]
    real*8 :: v0(3) = [4d5, 0d0, 0d0]
    real*8 :: B0(3) = [0d0, 0d0, 5d-9]
    logical :: gyration(8) = .true., fluid(8) = .false.
    ! Haser parameters (picard.f90:255-300: an H2O+ source as species 3; index
    ! 1 is never set there and reads zero); this host writes them to
    ! save/parameters.dat and does not run the source
    real*8 :: Qn(8) = [0d0, 0d0, (400.0d0)**(-3) * 1.99d26, 0d0, 0d0, 0d0, 0d0, 0d0]
    real*8 :: vn(8) = [0d0, 0d0, 629d0, 0d0, 0d0, 0d0, 0d0, 0d0]
    real*8 :: Ek(8) = [0d0, 0d0, 1.52d1 * EC, 0d0, 0d0, 0d0, 0d0, 0d0]
    real*8 :: nu_i(8) = [0d0, 0d0, 1.26d-7, 0d0, 0d0, 0d0, 0d0, 0d0]
    real*8 :: nu_d(8) = [0d0, 0d0, 1.26d-7, 0d0, 0d0, 0d0, 0d0, 0d0]
    ! output (picard.f90 section 1.1): dumps when dmod(time, interval) < dt,
    ! the diagnostics line every write_output_interval steps
    logical :: write_fields = .false., write_particles = .false.
    real*8 :: write_fields_interval = 1d30, write_particles_interval = 1d30
    integer :: write_output_interval = 0
  end type run_config

  type derived
    real*8 :: dxyz(3), dV, dt, omega_pe
    real*8 :: charge(8), density(8), n2p(8), eta(8), mass(8), kelvin(8)
    integer :: ppc(8), ipp(8), epp(8), hpp(8)
    real*8 :: v0(4,8), ve(4), B0(4), E0(4), tB(4,8), sB(4,8)
  end type derived

contains

  subroutine read_config(path, cfg)
    character(len=*), intent(in) :: path
    type(run_config), intent(inout) :: cfg
    integer :: nx, ny, nz, species, steps, maxit, seed, ppc(8), u, ios
    logical :: xflow, yflow, zflow, gyration(8), fluid(8)
    real*8 :: xmin, xmax, ymin, ymax, zmin, zmax, dt_wpe, maxerr, mass(8), density(8), kelvin(8), &
              v0(3), B0(3), write_fields_interval, write_particles_interval, Qn(8), vn(8), Ek(8), &
              nu_i(8), nu_d(8)
    logical :: write_fields, write_particles
    integer :: write_output_interval
    namelist /picard_run/ nx, ny, nz, species, steps, maxit, seed, ppc, xflow, yflow, zflow, gyration, &
                          fluid, xmin, xmax, ymin, ymax, zmin, zmax, dt_wpe, maxerr, mass, density, &
                          kelvin, v0, B0, write_fields, write_particles, write_fields_interval, &
                          write_particles_interval, write_output_interval, Qn, vn, Ek, nu_i, nu_d
    nx = cfg%nx; ny = cfg%ny; nz = cfg%nz; species = cfg%species; steps = cfg%steps
    maxit = cfg%maxit; seed = cfg%seed; ppc = cfg%ppc; xflow = cfg%xflow; yflow = cfg%yflow
    zflow = cfg%zflow; gyration = cfg%gyration; fluid = cfg%fluid; xmin = cfg%xmin; xmax = cfg%xmax
    ymin = cfg%ymin; ymax = cfg%ymax; zmin = cfg%zmin; zmax = cfg%zmax; dt_wpe = cfg%dt_wpe
    maxerr = cfg%maxerr; mass = cfg%mass; density = cfg%density; kelvin = cfg%kelvin; v0 = cfg%v0
    B0 = cfg%B0
    Qn = cfg%Qn; vn = cfg%vn; Ek = cfg%Ek; nu_i = cfg%nu_i; nu_d = cfg%nu_d
    write_fields = cfg%write_fields; write_particles = cfg%write_particles
    write_fields_interval = cfg%write_fields_interval
    write_particles_interval = cfg%write_particles_interval
    write_output_interval = cfg%write_output_interval
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'picard: cannot open ', trim(path)
      error stop 1
    end if
    read(u, nml=picard_run)
    close(u)
    cfg%nx = nx; cfg%ny = ny; cfg%nz = nz; cfg%species = species; cfg%steps = steps
    cfg%maxit = maxit; cfg%seed = seed; cfg%ppc = ppc; cfg%xflow = xflow; cfg%yflow = yflow
    cfg%zflow = zflow; cfg%gyration = gyration; cfg%fluid = fluid; cfg%xmin = xmin; cfg%xmax = xmax
    cfg%ymin = ymin; cfg%ymax = ymax; cfg%zmin = zmin; cfg%zmax = zmax; cfg%dt_wpe = dt_wpe
    cfg%maxerr = maxerr; cfg%mass = mass; cfg%density = density; cfg%kelvin = kelvin; cfg%v0 = v0
    cfg%B0 = B0
    cfg%Qn = Qn; cfg%vn = vn; cfg%Ek = Ek; cfg%nu_i = nu_i; cfg%nu_d = nu_d
    cfg%write_fields = write_fields; cfg%write_particles = write_particles
    cfg%write_fields_interval = write_fields_interval
    cfg%write_particles_interval = write_particles_interval
    cfg%write_output_interval = write_output_interval
  end subroutine read_config

  ! picard.f90:333-466 on one domain (iprocs = jprocs = kprocs = 1).
  ! u0 = uE0 + uB0 + uP0 of picard.f90:476-480, the scale of the diagnostics line
  real*8 function initial_energy(cfg, d) result(u0)
    type(run_config), intent(in) :: cfg
    type(derived), intent(in) :: d
    real*8 :: vol, uE0, uB0, uP0
    integer :: s
    vol = (cfg%xmax - cfg%xmin) * (cfg%ymax - cfg%ymin) * (cfg%zmax - cfg%zmin)
    uE0 = 0.5d0 * (d%E0(4)**2 * EPS0) * vol
    uB0 = 0.5d0 * (d%B0(4)**2 / MU0) * vol
    uP0 = 0d0
    do s = 1, 8
      uP0 = uP0 + d%density(s) * (d%mass(s) * d%v0(4, s)**2 + 3d0 * KB * d%kelvin(s))
    end do
    uP0 = 0.5d0 * uP0 * vol
    u0 = uE0 + uB0 + uP0
  end function initial_energy

  subroutine derive(cfg, d)
    type(run_config), intent(in) :: cfg
    type(derived), intent(out) :: d
    integer :: s
    real*8 :: w(8), den
    d%dxyz(1) = (cfg%xmax - cfg%xmin) / dble(cfg%nx)
    d%dxyz(2) = (cfg%ymax - cfg%ymin) / dble(cfg%ny)
    d%dxyz(3) = (cfg%zmax - cfg%zmin) / dble(cfg%nz)
    d%dV = d%dxyz(1) * d%dxyz(2) * d%dxyz(3)
    d%mass = cfg%mass
    d%kelvin = cfg%kelvin
    d%charge = EC
    d%charge(1) = -EC
    d%ppc = cfg%ppc
    d%density = cfg%density
    d%density(2:) = d%density(2:) * dble(d%ppc(2:)) / dble(max(d%ppc(2:), 1))
    d%density(1) = sum(-d%density(2:) * d%charge(2:)) * d%charge(1)**(-1)
    do s = 1, 8
      if (d%density(s) <= 0d0 .or. cfg%species < s) then
        d%ppc(s) = 0
        d%density(s) = 0d0
      end if
    end do
    d%density(1) = sum(-d%density(2:) * d%charge(2:) / d%charge(1))
    d%epp = max(int(0.5d0 - d%charge / d%charge(1)), 0)
    d%hpp = max(int(0.5d0 - d%charge / d%charge(2)), 0)
    d%hpp(1) = 0
    d%ipp(3:) = d%ppc(3:)
    d%ipp(2) = d%ppc(2) - sum(d%ipp(3:) * d%hpp(3:))
    d%ipp(1) = 0
    d%ppc(1) = sum(d%ipp(2:) * d%epp(2:))
    d%n2p = 0d0
    do s = 1, 8
      if (d%ppc(s) > 0) then
        d%n2p(s) = d%dV * d%density(s) * dble(d%ppc(s))**(-1)
      else
        d%n2p(s) = d%n2p(1)
      end if
    end do
    d%eta = d%charge / d%mass
    do s = 1, 8
      d%v0(1:3, s) = cfg%v0
    end do
    w = d%ipp * (d%n2p * d%mass + d%epp * d%n2p(1) * d%mass(1) + d%hpp * d%n2p(2) * d%mass(2))
    den = sum(w)
    d%ve = 0d0
    if (den /= 0d0) then
      d%ve(1) = sum(w * d%v0(1, :)) / den
      d%ve(2) = sum(w * d%v0(2, :)) / den
      d%ve(3) = sum(w * d%v0(3, :)) / den
    end if
    do s = 1, 8
      d%v0(4, s) = sqrt(sum(d%v0(1:3, s)**2))
    end do
    d%ve(4) = sqrt(sum(d%ve(1:3)**2))
    d%B0(1:3) = cfg%B0
    d%B0(4) = sqrt(sum(d%B0(1:3)**2))
    d%E0(1) = d%ve(3) * d%B0(2) - d%ve(2) * d%B0(3)
    d%E0(2) = d%ve(1) * d%B0(3) - d%ve(3) * d%B0(1)
    d%E0(3) = d%ve(2) * d%B0(1) - d%ve(1) * d%B0(2)
    d%E0(4) = sqrt(sum(d%E0(1:3)**2))
    d%omega_pe = sqrt(d%density(1) * EC * EC / (EPS0 * d%mass(1)))
    d%dt = cfg%dt_wpe / d%omega_pe
    do s = 1, 8
      if (cfg%gyration(s)) then
        d%tB(:, s) = 0.5d0 * d%dt * d%B0(:) * d%eta(s)
      else
        d%tB(:, s) = 0d0
      end if
    end do
    do s = 1, 4
      d%sB(s, :) = 2d0 * d%tB(s, :) / (1d0 + d%tB(4, :)**2)
    end do
  end subroutine derive

  subroutine seed_rng(seed)
    integer, intent(in) :: seed
    integer :: n, i
    integer, allocatable :: s(:)
    call random_seed(size=n)
    allocate(s(n))
    do i = 1, n
      s(i) = seed + 7919 * i
    end do
    call random_seed(put=s)
  end subroutine seed_rng

  ! Box-Muller, Tools.f90:30-48
  real*8 function normal_draw()
    real*8 :: r1, r2
    call random_number(r1)
    do
      call random_number(r2)
      if (r2 > 0d0) exit
    end do
    normal_draw = sqrt(-2d0 * log(r2)) * sin(TAU * r1)
  end function normal_draw

  ! Drifting Maxwellian, Tools.f90:2-28
  subroutine maxwellian(u, v, t, m)
    real*8, intent(out) :: u(3)
    real*8, intent(in) :: v(3), t, m
    real*8 :: a, pw
    integer :: c
    if (t > 0d0 .and. m > 0d0) then
      a = m / (2d0 * KB * t)
      pw = sqrt(a**(-1))
      do c = 1, 3
        u(c) = normal_draw()
      end do
      u = u * pw / sqrt(2d0)
      u = u + v
    else
      u = v
    end if
  end subroutine maxwellian

  ! Velocity half-step rotation with norm restore (ParticlesInitialize.f90:73-80)
  subroutine half_rotate(v, ve, tB)
    real*8, intent(inout) :: v(3)
    real*8, intent(in) :: ve(4), tB(4)
    real*8 :: vmin(3), vmid(3)
    vmin = v - ve(1:3)
    vmid(1) = vmin(1) + vmin(2) * tB(3) - vmin(3) * tB(2)
    vmid(2) = vmin(2) + vmin(3) * tB(1) - vmin(1) * tB(3)
    vmid(3) = vmin(3) + vmin(1) * tB(2) - vmin(2) * tB(1)
    if (sum(vmid**2) > 0d0) vmid = vmid * sqrt(sum(vmin**2) / sum(vmid**2))
    v = vmid + ve(1:3)
  end subroutine half_rotate

  ! One cell's plasma: each ion followed by its electrons and protons.
  subroutine emit_cell(d, species, first_species, x0, rp, sp, n, nmax)
    type(derived), intent(in) :: d
    integer, intent(in) :: species, first_species, nmax
    real*8, intent(in) :: x0(3)
    real*8, intent(inout) :: rp(8, nmax)
    integer, intent(inout) :: sp(nmax), n
    integer :: ss, pp, ee
    real*8 :: r(3)
    do ss = first_species, species
      do pp = 1, d%ipp(ss)
        if (n + 1 + d%epp(ss) + d%hpp(ss) > nmax) then
          write(*,*) 'picard: particle capacity exceeded by the sources'
          error stop 1
        end if
        call random_number(r)
        rp(:, n + 1) = 0d0
        rp(1:3, n + 1) = x0 + d%dxyz * r
        sp(n + 1) = ss
        call maxwellian(rp(4:6, n + 1), d%v0(1:3, ss), d%kelvin(ss), d%mass(ss))
        call half_rotate(rp(4:6, n + 1), d%ve, d%tB(:, ss))
        do ee = 1, d%epp(ss)
          rp(:, n + 1 + ee) = rp(:, n + 1)
          sp(n + 1 + ee) = 1
          call maxwellian(rp(4:6, n + 1 + ee), d%v0(1:3, ss), d%kelvin(1), d%mass(1))
          call half_rotate(rp(4:6, n + 1 + ee), d%ve, d%tB(:, 1))
        end do
        do ee = 1, d%hpp(ss)
          rp(:, n + 1 + d%epp(ss) + ee) = rp(:, n + 1)
          sp(n + 1 + d%epp(ss) + ee) = 2
          call maxwellian(rp(4:6, n + 1 + d%epp(ss) + ee), d%v0(1:3, ss), d%kelvin(2), d%mass(2))
          call half_rotate(rp(4:6, n + 1 + d%epp(ss) + ee), d%ve, d%tB(:, 2))
        end do
        n = n + 1 + d%epp(ss) + d%hpp(ss)
      end do
    end do
  end subroutine emit_cell

  ! ParticlesInitialize.f90:54-119 on one domain
  subroutine initialize_particles(cfg, d, rp, sp, n, nmax)
    type(run_config), intent(in) :: cfg
    type(derived), intent(in) :: d
    integer, intent(in) :: nmax
    real*8, intent(inout) :: rp(8, nmax)
    integer, intent(inout) :: sp(nmax), n
    integer :: i, j, k
    real*8 :: lo(3)
    lo = [cfg%xmin, cfg%ymin, cfg%zmin]
    do k = 2, cfg%nz + 1
      do j = 2, cfg%ny + 1
        do i = 2, cfg%nx + 1
          call emit_cell(d, cfg%species, 1, lo + d%dxyz * dble([i, j, k] - 2), rp, sp, n, nmax)
        end do
      end do
    end do
  end subroutine initialize_particles

  ! BCinflow.f90: fresh plasma in the guard cells of every flow face
  subroutine inflow_particles(cfg, d, rp, sp, n, nmax)
    type(run_config), intent(in) :: cfg
    type(derived), intent(in) :: d
    integer, intent(in) :: nmax
    real*8, intent(inout) :: rp(8, nmax)
    integer, intent(inout) :: sp(nmax), n
    integer :: i, j, k, yn, yp, zn, zp
    real*8 :: lo(3)
    lo = [cfg%xmin, cfg%ymin, cfg%zmin]
    yn = merge(1, 0, cfg%yflow); yp = yn
    zn = merge(1, 0, cfg%zflow); zp = zn
    if (cfg%xflow) then
      do k = 2 - zn, cfg%nz + 1 + zp
        do j = 2 - yn, cfg%ny + 1 + yp
          call emit_cell(d, cfg%species, 1, lo + d%dxyz * dble([cfg%nx + 2, j, k] - 2), rp, sp, n, nmax)
        end do
      end do
      do k = 2 - zn, cfg%nz + 1 + zp
        do j = 2 - yn, cfg%ny + 1 + yp
          call emit_cell(d, cfg%species, 1, lo + d%dxyz * dble([1, j, k] - 2), rp, sp, n, nmax)
        end do
      end do
    end if
    if (cfg%yflow) then
      do k = 2, cfg%nz + 1 + zp
        do i = 2, cfg%nx + 1
          call emit_cell(d, cfg%species, 2, lo + d%dxyz * dble([i, cfg%ny + 2, k] - 2), rp, sp, n, nmax)
        end do
      end do
      do k = 2 - zn, cfg%nz + 1
        do i = 2, cfg%nx + 1
          call emit_cell(d, cfg%species, 2, lo + d%dxyz * dble([i, 1, k] - 2), rp, sp, n, nmax)
        end do
      end do
    end if
    if (cfg%zflow) then
      do j = 2, cfg%ny + 1
        do i = 2, cfg%nx + 1
          call emit_cell(d, cfg%species, 2, lo + d%dxyz * dble([i, j, cfg%nz + 2] - 2), rp, sp, n, nmax)
        end do
      end do
      do j = 2, cfg%ny + 1
        do i = 2, cfg%nx + 1
          call emit_cell(d, cfg%species, 2, lo + d%dxyz * dble([i, j, 1] - 2), rp, sp, n, nmax)
        end do
      end do
    end if
  end subroutine inflow_particles

  subroutine fill_structs(cfg, d, max_particles, geo, spc)
    type(run_config), intent(in) :: cfg
    type(derived), intent(in) :: d
    integer, intent(in) :: max_particles
    type(picard_geometry), intent(out) :: geo
    type(picard_species), intent(out) :: spc
    integer :: s
    geo%nx = cfg%nx; geo%ny = cfg%ny; geo%nz = cfg%nz; geo%species = cfg%species
    geo%flow = [merge(1, 0, cfg%xflow), merge(1, 0, cfg%yflow), merge(1, 0, cfg%zflow)]
    geo%precision = PICARD_FP64
    geo%deposit = PICARD_DEPOSIT_REFERENCE_ORDER
    geo%device = 0
    geo%max_particles = max_particles
    geo%gmin = [cfg%xmin, cfg%ymin, cfg%zmin]
    geo%gmax = [cfg%xmax, cfg%ymax, cfg%zmax]
    geo%lmin = geo%gmin
    geo%lmax = geo%gmax
    geo%dxyz = d%dxyz
    spc%charge = d%charge; spc%n2p = d%n2p; spc%eta = d%eta
    do s = 1, 8
      spc%tB(4*(s-1)+1:4*s) = d%tB(:, s)
      spc%sB(4*(s-1)+1:4*s) = d%sB(:, s)
    end do
    spc%ve = d%ve; spc%B0 = d%B0; spc%E0 = d%E0
    spc%fluid = merge(1, 0, cfg%fluid)
    spc%gyration = merge(1, 0, cfg%gyration)
  end subroutine fill_structs

end module picard_setup
