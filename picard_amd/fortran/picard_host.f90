! picard_host.f90 -- Fortran host of the PIC time step on one MI355X.
!
! The main loop keeps the reference's order and call surface
! (picard.f90:554-633); the subroutines it calls are the drop-in versions in
! picard_dropin.f90, which run on the GPU through libpicard_hip.so.
!
!   picard_host [config.nml]      (namelist &picard_run, see picard_setup.f90)
program picard_host
  use iso_c_binding
  use picard_hip_iface
  use picard_dropin_state
  use picard_setup
  use picard_output
  implicit none

  type(run_config) :: cfg
  type(derived) :: dc
  type(picard_geometry) :: geo
  type(picard_species) :: spc
  real*8, allocatable :: real_particles(:,:), U(:,:,:), E(:,:,:,:), N(:,:,:,:), C(:,:,:), D(:,:,:)
  real*8, allocatable :: F(:,:,:,:,:)
  real*8 :: time, uP, uE, uB, u0
  integer :: num_global, fields_dump, particles_dump
  logical :: out_fields, out_particles
  integer, allocatable :: int_particles(:)
  integer :: max_per_proc, num_local, iteration, nargs, per_cell
  integer(c_int64_t) :: pushed, sweeps_total
  integer(8) :: t0, t1, rate
  character(len=512) :: path
  real*8 :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, secs

  nargs = command_argument_count()
  if (nargs >= 1) then
    call get_command_argument(1, path)
    call read_config(trim(path), cfg)
  end if
  call derive(cfg, dc)
  call seed_rng(cfg%seed)

  per_cell = sum(dc%ppc(1:cfg%species))
  max_per_proc = int(1.5d0 * per_cell * (cfg%nx + 2) * (cfg%ny + 2) * (cfg%nz + 2)) + 1024
  allocate(real_particles(8, max_per_proc), int_particles(max_per_proc))
  allocate(U(cfg%nx+2, cfg%ny+2, cfg%nz+2), C(cfg%nx+2, cfg%ny+2, cfg%nz+2))
  allocate(E(3, cfg%nx+2, cfg%ny+2, cfg%nz+2), N(cfg%species, cfg%nx+2, cfg%ny+2, cfg%nz+2))
  allocate(D(2*cfg%nx+1, 2*cfg%ny+1, 2*cfg%nz+1))
  allocate(F(cfg%species, 3, cfg%nx+2, cfg%ny+2, cfg%nz+2))
  F = 0d0
  real_particles = 0d0; int_particles = 0; U = 0d0; E = 0d0; N = 0d0; C = 0d0; D = 0d0
  Lx_min = cfg%xmin; Lx_max = cfg%xmax; Ly_min = cfg%ymin; Ly_max = cfg%ymax
  Lz_min = cfg%zmin; Lz_max = cfg%zmax

  ! the run's parameters, as picard.f90:488-489 writes them at setup
  call dump_parameters(dc%n2p, dc%mass, dc%charge, dc%density, dc%kelvin, cfg%Qn, cfg%vn, &
                       cfg%Ek, cfg%nu_i, cfg%nu_d, dc%ppc, cfg%gyration, dc%ve, dc%B0, dc%dxyz, &
                       [cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax], dc%dt, &
                       [cfg%nx, cfg%ny, cfg%nz], [1, 1, 1], [cfg%xflow, cfg%yflow, cfg%zflow])
  call fill_structs(cfg, dc, max_per_proc, geo, spc)
  call picard_dropin_init(geo, spc, real_particles, int_particles, max_per_proc)

  num_local = 0
  call initialize_particles(cfg, dc, real_particles, int_particles, num_local, max_per_proc)
  call CalcDistance(D, cfg%nx, cfg%ny, cfg%nz, dc%dxyz, cfg%xflow, cfg%yflow, cfg%zflow)
  write(*,'(a,3i5,a,i0,a,es12.5)') 'picard_host: grid', cfg%nx, cfg%ny, cfg%nz, '  particles ', &
      num_local, '  dt ', dc%dt

  ! the initial energy the diagnostics line is divided by (picard.f90:476-480)
  u0 = initial_energy(cfg, dc)
  time = 0d0
  fields_dump = 0
  particles_dump = 0
  pushed = 0
  sweeps_total = 0
  call system_clock(t0, rate)
  do iteration = 1, cfg%steps
    time = time + dc%dt                                                                       ! :556
    call inflow_particles(cfg, dc, real_particles, int_particles, num_local, max_per_proc)   ! :560
    call AdvancePosition(real_particles, dc%dt, max_per_proc, num_local)                      ! :564
    call BCoutflow(real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                   cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, max_per_proc, num_local, &
                   cfg%xflow, cfg%yflow, cfg%zflow)                                          ! :569
    call ParticlesTransfer(real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                           cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, max_per_proc, &
                           num_local, 1, 1, 1, 0)                                            ! :572
    call ParticlesGridify(N, C, real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, &
                          Lz_max, dc%dxyz, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, &
                          cfg%species, cfg%nx, cfg%ny, cfg%nz, max_per_proc, num_local, cfg%xflow, &
                          cfg%yflow, cfg%zflow, 1, 1, 1, dc%charge, dc%n2p, 0)                 ! :577
    call BCpotentials(U, C, D, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, cfg%zmax, cfg%nx, cfg%ny, &
                      cfg%nz, cfg%nx, cfg%ny, cfg%nz, cfg%xflow, cfg%yflow, cfg%zflow, cfg%species, &
                      1, 1, 1, 0)                                                            ! :581
    call CalcEpotential(U, C, cfg%nx, cfg%ny, cfg%nz, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                        cfg%xflow, cfg%yflow, cfg%zflow, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, &
                        cfg%zmax, dc%dxyz, cfg%maxerr, cfg%maxit, 1, 1, 1, cfg%species, 0)    ! :584
    call CalcEfield(E, U, cfg%nx, cfg%ny, cfg%nz, dc%dxyz, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, 0)  ! :587
    call BCfields(E, U, dc%B0, dc%E0, cfg%nx, cfg%ny, cfg%nz, cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, 0) ! :589
    ! output steps split the velocity push around the dumps (:600-628)
    out_fields = cfg%write_fields .and. dmod(time, cfg%write_fields_interval) .lt. dc%dt
    out_particles = cfg%write_particles .and. dmod(time, cfg%write_particles_interval) .lt. dc%dt
    if (out_fields .or. out_particles) then
      call AdvanceVelocityImp(real_particles, int_particles, E, dc%B0, Lx_min, Ly_min, Lz_min, dc%dxyz, &
                              0.5d0*dc%dt, dc%dt, dc%v0, dc%ve, dc%eta, dc%tB, num_local, max_per_proc, &
                              cfg%species, cfg%nx, cfg%ny, cfg%nz, cfg%fluid, cfg%gyration)  ! :603
      if (out_particles) then
        call picard_dropin_download_particles()
        call dump_particles(real_particles, int_particles, num_local, particles_dump, 0)      ! :607
        particles_dump = particles_dump + 1
      end if
      if (out_fields) then
        call ParticlesGridifyAll(F, N, C, real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, &
                                 Lz_min, Lz_max, dc%dxyz, cfg%xmin, cfg%xmax, cfg%ymin, cfg%ymax, cfg%zmin, &
                                 cfg%zmax, cfg%species, cfg%nx, cfg%ny, cfg%nz, max_per_proc, num_local, &
                                 cfg%xflow, cfg%yflow, cfg%zflow, 1, 1, 1, dc%charge, dc%n2p, 0)  ! :611
        call picard_dropin_download_field(PICARD_FIELD_U, U, size(U, kind=c_int64_t))
        call picard_dropin_download_field(PICARD_FIELD_E, E, size(E, kind=c_int64_t))
        call dump_fields(U, N, F, E, cfg%species, fields_dump, 0)                            ! :614
        fields_dump = fields_dump + 1
      end if
      call AdvanceVelocityExp(real_particles, int_particles, E, dc%B0, Lx_min, Ly_min, Lz_min, dc%dxyz, &
                              0.5d0*dc%dt, dc%dt, dc%v0, dc%ve, dc%eta, dc%tB, num_local, max_per_proc, &
                              cfg%species, cfg%nx, cfg%ny, cfg%nz, cfg%fluid, cfg%gyration)  ! :616
    else
      call AdvanceVelocity(real_particles, int_particles, E, dc%B0, Lx_min, Ly_min, Lz_min, dc%dxyz, dc%dt, &
                           dc%v0, dc%ve, dc%eta, dc%tB, dc%sB, num_local, max_per_proc, cfg%species, cfg%nx, &
                           cfg%ny, cfg%nz, cfg%fluid, cfg%gyration)                          ! :630
    end if
    if (cfg%write_output_interval > 0) then
      if (mod(iteration, cfg%write_output_interval) == 0) then                               ! :647-672
        call picard_dropin_diagnostics(dc%mass, num_global, uP, uE, uB)
        write(*, fmt='(I8,a,E12.6,a,I8,a,I8,a,I8,a,E12.6,a,E12.6,a,E12.6)') iteration, '   ', time, &
            '   ', 0, '   ', num_local, '   ', num_global, '   ', uP/u0, '   ', uE/u0, '   ', uB/u0
      end if
    end if
    pushed = pushed + num_local
    sweeps_total = sweeps_total + last_sweeps
    write(*,'(a,i6,a,i9,a,i7,a,i8)') 'step', iteration, '  num_local', num_local, '  removed', &
        last_removed, '  sweeps', last_sweeps
  end do
  call system_clock(t1)
  secs = dble(t1 - t0) / dble(rate)
  call picard_dropin_download_particles()
  write(*,'(a,es12.5,a,f10.3,a)') 'picard_host: particle-steps/s ', dble(pushed) / secs, ' over ', secs, ' s'
  write(*,'(a,es24.16)') 'picard_host: checksum sum(|v|) ', sum(abs(real_particles(4:6, 1:num_local)))
  call picard_dropin_finalize()
end program picard_host
