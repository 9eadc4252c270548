! dropin_check.f90 -- the Fortran drop-in subroutines (GPU) against the C
! oracle on the same host-generated particles, bit for bit, step by step.
!   dropin_check [steps]      prints 'DROPIN OK' and exits 0 on parity
!   dropin_check mismatch K   calls one drop-in with argument K changed from
!                             the domain given at init (xflow, dxyz, tB,
!                             max_per_proc, n2p, E0, myid); the call must stop
!                             with an error before any device work
program dropin_check
  use iso_c_binding
  use picard_hip_iface
  use picard_dropin_state
  use picard_setup
  implicit none

  type, bind(C) :: orc_grid
    integer(c_int32_t) :: nx, ny, nz, species
    integer(c_int32_t) :: flow(3)
    real(c_double) :: lmin(3), lmax(3), gmin(3), gmax(3), dxyz(3)
  end type orc_grid

  interface
    subroutine orc_advance_position(rp, dt, n) bind(C, name='orc_advance_position')
      import :: c_double, c_int64_t
      real(c_double) :: rp(*)
      real(c_double), value :: dt
      integer(c_int64_t), value :: n
    end subroutine
    integer(c_int64_t) function orc_bc_outflow(g, rp, sp, n, removed) bind(C, name='orc_bc_outflow')
      import :: c_double, c_int64_t, c_int32_t, c_ptr, orc_grid
      type(orc_grid) :: g
      real(c_double) :: rp(*)
      integer(c_int32_t) :: sp(*)
      integer(c_int64_t) :: n
      type(c_ptr), value :: removed
    end function
    subroutine orc_particles_transfer(g, rp, n) bind(C, name='orc_particles_transfer')
      import :: c_double, c_int64_t, orc_grid
      type(orc_grid) :: g
      real(c_double) :: rp(*)
      integer(c_int64_t), value :: n
    end subroutine
    integer(c_int64_t) function orc_particles_gridify(g, Ng, Cg, rp, sp, n, charge, n2p) &
        bind(C, name='orc_particles_gridify')
      import :: c_double, c_int64_t, c_int32_t, orc_grid
      type(orc_grid) :: g
      real(c_double) :: Ng(*), Cg(*), rp(*), charge(8), n2p(8)
      integer(c_int32_t) :: sp(*)
      integer(c_int64_t), value :: n
    end function
    subroutine orc_calc_distance(D, Nx, Ny, Nz, dxyz, flow) bind(C, name='orc_calc_distance')
      import :: c_double, c_int32_t
      real(c_double) :: D(*), dxyz(3)
      integer(c_int32_t), value :: Nx, Ny, Nz
      integer(c_int32_t) :: flow(3)
    end subroutine
    subroutine orc_bc_potentials(g, U, C, D) bind(C, name='orc_bc_potentials')
      import :: c_double, orc_grid
      type(orc_grid) :: g
      real(c_double) :: U(*), C(*), D(*)
    end subroutine
    integer(c_int64_t) function orc_calc_epotential(g, U, C, maxerr, maxit, ratio) &
        bind(C, name='orc_calc_epotential')
      import :: c_double, c_int64_t, c_int32_t, c_ptr, orc_grid
      type(orc_grid) :: g
      real(c_double) :: U(*), C(*)
      real(c_double), value :: maxerr
      integer(c_int32_t), value :: maxit
      type(c_ptr), value :: ratio
    end function
    subroutine orc_calc_efield(g, E, U) bind(C, name='orc_calc_efield')
      import :: c_double, orc_grid
      type(orc_grid) :: g
      real(c_double) :: E(*), U(*)
    end subroutine
    subroutine orc_bc_fields(g, E) bind(C, name='orc_bc_fields')
      import :: c_double, orc_grid
      type(orc_grid) :: g
      real(c_double) :: E(*)
    end subroutine
    subroutine orc_advance_velocity(g, rp, sp, E, B0, dt, ve, eta, tB, sB, n, fluid, gyration) &
        bind(C, name='orc_advance_velocity')
      import :: c_double, c_int64_t, c_int32_t, orc_grid
      type(orc_grid) :: g
      real(c_double) :: rp(*), E(*), B0(4), ve(4), eta(8), tB(32), sB(32)
      integer(c_int32_t) :: sp(*), fluid(8), gyration(8)
      real(c_double), value :: dt
      integer(c_int64_t), value :: n
    end subroutine
  end interface

  type(run_config) :: cfg
  type(derived) :: dc
  type(picard_geometry) :: geo
  type(picard_species) :: spc
  type(orc_grid) :: og
  real*8, allocatable :: rp(:,:), U(:,:,:), E(:,:,:,:), N(:,:,:,:), C(:,:,:), Dt(:,:,:)
  real*8, allocatable :: rp_o(:,:), U_o(:,:,:), E_o(:,:,:,:), N_o(:,:,:,:), C_o(:,:,:), D_o(:,:,:)
  integer, allocatable :: sp(:), sp_o(:)
  integer :: maxp, nl, it, steps, first, nargs, bad
  integer(c_int64_t) :: cnt_o, sw_o, dummy
  character(len=32) :: arg, mism
  real*8 :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max

  steps = 4
  mism = ''
  nargs = command_argument_count()
  if (nargs >= 1) then
    call get_command_argument(1, arg)
    if (arg == 'mismatch') then
      call get_command_argument(2, mism)
    else
      read(arg, *) steps
    end if
  end if
  cfg%nx = 10; cfg%ny = 9; cfg%nz = 8; cfg%ppc(2) = 4
  cfg%xmin = -50d0; cfg%xmax = 50d0; cfg%ymin = -45d0; cfg%ymax = 45d0; cfg%zmin = -40d0; cfg%zmax = 40d0
  cfg%yflow = .true.
  call derive(cfg, dc)
  call seed_rng(cfg%seed)
  maxp = 40 * (cfg%nx + 2) * (cfg%ny + 2) * (cfg%nz + 2)
  allocate(rp(8, maxp), sp(maxp), rp_o(8, maxp), sp_o(maxp))
  allocate(U(cfg%nx+2, cfg%ny+2, cfg%nz+2), C(cfg%nx+2, cfg%ny+2, cfg%nz+2), U_o(cfg%nx+2, cfg%ny+2, cfg%nz+2))
  allocate(C_o(cfg%nx+2, cfg%ny+2, cfg%nz+2), E(3, cfg%nx+2, cfg%ny+2, cfg%nz+2), E_o(3, cfg%nx+2, cfg%ny+2, cfg%nz+2))
  allocate(N(cfg%species, cfg%nx+2, cfg%ny+2, cfg%nz+2), N_o(cfg%species, cfg%nx+2, cfg%ny+2, cfg%nz+2))
  allocate(Dt(2*cfg%nx+1, 2*cfg%ny+1, 2*cfg%nz+1), D_o(2*cfg%nx+1, 2*cfg%ny+1, 2*cfg%nz+1))
  rp = 0d0; sp = 0; U = 0d0; E = 0d0; N = 0d0; C = 0d0; Dt = 0d0
  Lx_min = cfg%xmin; Lx_max = cfg%xmax; Ly_min = cfg%ymin; Ly_max = cfg%ymax; Lz_min = cfg%zmin; Lz_max = cfg%zmax

  og%nx = cfg%nx; og%ny = cfg%ny; og%nz = cfg%nz; og%species = cfg%species
  og%flow = [merge(1, 0, cfg%xflow), merge(1, 0, cfg%yflow), merge(1, 0, cfg%zflow)]
  og%gmin = [cfg%xmin, cfg%ymin, cfg%zmin]; og%gmax = [cfg%xmax, cfg%ymax, cfg%zmax]
  og%lmin = og%gmin; og%lmax = og%gmax; og%dxyz = dc%dxyz

  call fill_structs(cfg, dc, maxp, geo, spc)
  call picard_dropin_init(geo, spc, rp, sp, maxp)
  eager_download = .true.

  nl = 0
  call initialize_particles(cfg, dc, rp, sp, nl, maxp)
  rp_o(:, 1:nl) = rp(:, 1:nl); sp_o(1:nl) = sp(1:nl); cnt_o = nl
  call CalcDistance(Dt, cfg%nx, cfg%ny, cfg%nz, dc%dxyz, cfg%xflow, cfg%yflow, cfg%zflow)
  call orc_calc_distance(D_o, cfg%nx, cfg%ny, cfg%nz, dc%dxyz, og%flow)
  bad = 0
  call cmp('D', Dt, D_o, size(Dt))
  if (mism /= '') call mismatched_call(trim(mism))

  do it = 1, steps
    first = nl + 1
    call inflow_particles(cfg, dc, rp, sp, nl, maxp)
    rp_o(:, first:nl) = rp(:, first:nl); sp_o(first:nl) = sp(first:nl); cnt_o = nl
    ! GPU through the drop-in subroutines
    call AdvancePosition(rp, dc%dt, maxp, nl)
    call BCoutflow(rp, sp, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, cfg%xmin, cfg%xmax, cfg%ymin, &
                   cfg%ymax, cfg%zmin, cfg%zmax, maxp, nl, cfg%xflow, cfg%yflow, cfg%zflow)
    call ParticlesTransfer(rp, sp, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, cfg%xmin, cfg%xmax, &
                           cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, maxp, nl, 1, 1, 1, 0)
    call ParticlesGridify(N, C, rp, sp, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, dc%dxyz, cfg%xmin, &
                          cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, cfg%species, cfg%nx, cfg%ny, &
                          cfg%nz, maxp, nl, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, dc%charge, dc%n2p, 0)
    call BCpotentials(U, C, Dt, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, cfg%nx, cfg%ny, &
                      cfg%nz, cfg%nx, cfg%ny, cfg%nz, cfg%xflow, cfg%yflow, cfg%zflow, cfg%species, 1, 1, 1, 0)
    call CalcEpotential(U, C, cfg%nx, cfg%ny, cfg%nz, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                        cfg%xflow, cfg%yflow, cfg%zflow, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, &
                        cfg%zmax, dc%dxyz, cfg%maxerr, cfg%maxit, 1, 1, 1, cfg%species, 0)
    call CalcEfield(E, U, cfg%nx, cfg%ny, cfg%nz, dc%dxyz, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, 0)
    call BCfields(E, U, dc%B0, dc%E0, cfg%nx, cfg%ny, cfg%nz, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, 0)
    call AdvanceVelocity(rp, sp, E, dc%B0, Lx_min, Ly_min, Lz_min, dc%dxyz, dc%dt, dc%v0, dc%ve, dc%eta, &
                         dc%tB, dc%sB, nl, maxp, cfg%species, cfg%nx, cfg%ny, cfg%nz, cfg%fluid, cfg%gyration)
    ! oracle
    call orc_advance_position(rp_o, dc%dt, cnt_o)
    dummy = orc_bc_outflow(og, rp_o, sp_o, cnt_o, c_null_ptr)
    call orc_particles_transfer(og, rp_o, cnt_o)
    dummy = orc_particles_gridify(og, N_o, C_o, rp_o, sp_o, cnt_o, dc%charge, dc%n2p)
    call orc_bc_potentials(og, U_o, C_o, D_o)
    sw_o = orc_calc_epotential(og, U_o, C_o, cfg%maxerr, int(cfg%maxit, c_int32_t), c_null_ptr)
    call orc_calc_efield(og, E_o, U_o)
    call orc_bc_fields(og, E_o)
    call orc_advance_velocity(og, rp_o, sp_o, E_o, dc%B0, dc%dt, dc%ve, dc%eta, reshape(dc%tB, [32]), &
                              reshape(dc%sB, [32]), cnt_o, merge(1, 0, cfg%fluid), merge(1, 0, cfg%gyration))
    write(*,'(a,i3,a,i7,a,i7,a,i6,a,i6)') 'step', it, '  num_local gpu', nl, ' oracle', cnt_o, &
        '  sweeps gpu', last_sweeps, ' oracle', sw_o
    if (nl /= cnt_o .or. last_sweeps /= sw_o) bad = bad + 1
    call cmp('particles', rp(:, 1:nl), rp_o(:, 1:nl), 8 * nl)
    if (any(sp(1:nl) /= sp_o(1:nl))) then
      write(*,*) 'species differ'
      bad = bad + 1
    end if
    call cmp('N', N, N_o, size(N))
    call cmp('C', C, C_o, size(C))
    call cmp('U', U, U_o, size(U))
    call cmp('E', E, E_o, size(E))
  end do
  call picard_dropin_finalize()
  if (bad /= 0) then
    write(*,'(a,i0,a)') 'DROPIN FAILED (', bad, ' mismatches)'
    error stop 1
  end if
  write(*,'(a)') 'DROPIN OK'

contains

  ! One drop-in call with one argument changed; returns only if the call
  ! accepted it (then the program stops with a failure of its own).
  subroutine mismatched_call(which)
    character(len=*), intent(in) :: which
    real*8 :: dx2(3), tB2(4, 8), n2p2(8), E02(4)
    write(*,'(a,a)') 'mismatch case ', which
    dx2 = dc%dxyz
    tB2 = dc%tB
    n2p2 = dc%n2p
    E02 = dc%E0
    select case (which)
    case ('xflow')
      call BCoutflow(rp, sp, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, cfg%xmin, cfg%xmax, cfg%ymin, &
                     cfg%ymax, cfg%zmin, cfg%zmax, maxp, nl, .not. cfg%xflow, cfg%yflow, cfg%zflow)
    case ('dxyz')
      dx2(2) = 2d0 * dx2(2)
      call CalcEfield(E, U, cfg%nx, cfg%ny, cfg%nz, dx2, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, 0)
    case ('tB')
      tB2(3, 1) = tB2(3, 1) * (1d0 + epsilon(1d0))
      call AdvanceVelocity(rp, sp, E, dc%B0, Lx_min, Ly_min, Lz_min, dc%dxyz, dc%dt, dc%v0, dc%ve, dc%eta, &
                           tB2, dc%sB, nl, maxp, cfg%species, cfg%nx, cfg%ny, cfg%nz, cfg%fluid, cfg%gyration)
    case ('max_per_proc')
      call AdvancePosition(rp, dc%dt, maxp - 1, nl)
    case ('n2p')
      n2p2(2) = 0.5d0 * n2p2(2)
      call ParticlesGridify(N, C, rp, sp, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, dc%dxyz, cfg%xmin, &
                            cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, cfg%species, cfg%nx, cfg%ny, &
                            cfg%nz, maxp, nl, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, dc%charge, n2p2, 0)
    case ('E0')
      E02(1) = E02(1) + 1d0
      call BCfields(E, U, dc%B0, E02, cfg%nx, cfg%ny, cfg%nz, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, 0)
    case ('myid')
      call CalcEpotential(U, C, cfg%nx, cfg%ny, cfg%nz, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                          cfg%xflow, cfg%yflow, cfg%zflow, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, &
                          cfg%zmax, dc%dxyz, cfg%maxerr, cfg%maxit, 1, 1, 1, cfg%species, 1)
    case default
      write(*,'(a,a)') 'unknown mismatch case ', which
      error stop 3
    end select
    write(*,'(a)') 'MISMATCH ACCEPTED'
    error stop 4
  end subroutine mismatched_call

  subroutine cmp(what, a, b, n)
    character(len=*), intent(in) :: what
    integer, intent(in) :: n
    real*8, intent(in) :: a(n), b(n)
    integer :: i, nbad
    nbad = 0
    do i = 1, n
      if (transfer(a(i), 0_8) /= transfer(b(i), 0_8)) nbad = nbad + 1
    end do
    if (nbad > 0) then
      write(*,'(a,a,i0,a)') what, ': ', nbad, ' values differ'
      bad = bad + 1
    end if
  end subroutine cmp

end program dropin_check
