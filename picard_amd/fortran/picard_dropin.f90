! picard_dropin.f90 -- the reference's time-step subroutines, with their exact
! names and argument lists, running on the MI355X through libpicard_hip.so.
!
! A host written against the reference's call surface (picard.f90:564-631)
! links these in place of AdvancePosition.f90, BCoutflow.f90,
! ParticlesTransfer.f90, ParticlesGridify.f90, CalcDistance.f90,
! BCpotentials.f90, CalcEpotential.f90, CalcEfield.f90, BCfields.f90 and
! AdvanceVelocity.f90.
!
! Residency: the host calls picard_dropin_init once with its particle and
! field arrays (static in the main program, as in picard.f90:91-113).  The
! device keeps particles and fields between calls.  Particles the host appends
! after num_local (BCinflow, ParticlesProduce) are uploaded by the next call;
! host copies of device state are refreshed only by picard_dropin_download,
! or after every call when eager_download is set.
!
! Arguments: the domain (cell counts, flow flags, local and global bounds,
! cell sizes, species count, max_per_proc) and the species tables (charge,
! n2p, eta, ve, tB, sB, B0, E0, fluid, gyration) are fixed by
! picard_dropin_init.  Every subroutine compares the ones it is passed with
! them, bit for bit, and stops with an error naming the argument on a
! mismatch: a host that changes them between calls would otherwise get the
! initial values silently.  dt, deltat, maxerr and maxit are forwarded on
! every call.  The field arrays (N, C, U, E, D) are device-resident: host
! writes to them between calls are not uploaded, and their host copies are
! refreshed only as described above.
module picard_dropin_state
  use iso_c_binding
  use picard_hip_iface
  implicit none

  type(c_ptr), save :: ctx = c_null_ptr
  integer(c_int64_t), save :: dev_count = 0        ! particles resident on the device
  type(c_ptr), save :: host_rp = c_null_ptr        ! registered real_particles(8, max_per_proc)
  type(c_ptr), save :: host_sp = c_null_ptr        ! registered int_particles(max_per_proc)
  integer, save :: host_max = 0
  logical, save :: eager_download = .false.
  integer(c_int64_t), save :: last_sweeps = 0
  integer(c_int64_t), save :: last_removed = 0
  type(picard_geometry), save :: geo0              ! the domain given to picard_dropin_init
  type(picard_species), save :: spc0

contains

  subroutine arg_mismatch(where, what)
    character(len=*), intent(in) :: where, what
    write(*,'(a,a,a,a,a)') 'picard: ', where, ': argument ', what, &
        ' differs from the domain given to picard_dropin_init'
    error stop 2
  end subroutine arg_mismatch

  logical function differs(a, b, n)
    integer, intent(in) :: n
    real(c_double), intent(in) :: a(n), b(n)
    integer :: i
    differs = .false.
    do i = 1, n
      if (transfer(a(i), 0_c_int64_t) /= transfer(b(i), 0_c_int64_t)) differs = .true.
    end do
  end function differs

  subroutine check_flows(where, xflow, yflow, zflow)
    character(len=*), intent(in) :: where
    logical, intent(in) :: xflow, yflow, zflow
    if ((xflow .neqv. (geo0%flow(1) /= 0)) .or. (yflow .neqv. (geo0%flow(2) /= 0)) .or. &
        (zflow .neqv. (geo0%flow(3) /= 0))) call arg_mismatch(where, 'xflow/yflow/zflow')
  end subroutine check_flows

  subroutine check_cells(where, nx, ny, nz)
    character(len=*), intent(in) :: where
    integer, intent(in) :: nx, ny, nz
    if (nx /= geo0%nx .or. ny /= geo0%ny .or. nz /= geo0%nz) &
      call arg_mismatch(where, 'Nx_local/Ny_local/Nz_local')
  end subroutine check_cells

  subroutine check_local_box(where, lo, hi)
    character(len=*), intent(in) :: where
    real(c_double), intent(in) :: lo(3), hi(3)
    if (differs(lo, geo0%lmin, 3) .or. differs(hi, geo0%lmax, 3)) &
      call arg_mismatch(where, 'Lx_min..Lz_max')
  end subroutine check_local_box

  subroutine check_global_box(where, lo, hi)
    character(len=*), intent(in) :: where
    real(c_double), intent(in) :: lo(3), hi(3)
    if (differs(lo, geo0%gmin, 3) .or. differs(hi, geo0%gmax, 3)) &
      call arg_mismatch(where, 'xmin..zmax')
  end subroutine check_global_box

  subroutine check_dxyz(where, dxyz)
    character(len=*), intent(in) :: where
    real(c_double), intent(in) :: dxyz(3)
    if (differs(dxyz, geo0%dxyz, 3)) call arg_mismatch(where, 'dxyz')
  end subroutine check_dxyz

  subroutine check_species(where, species)
    character(len=*), intent(in) :: where
    integer, intent(in) :: species
    if (species /= geo0%species) call arg_mismatch(where, 'species')
  end subroutine check_species

  subroutine check_capacity(where, max_per_proc)
    character(len=*), intent(in) :: where
    integer, intent(in) :: max_per_proc
    if (max_per_proc /= host_max) call arg_mismatch(where, 'max_per_proc')
  end subroutine check_capacity

  subroutine check_charges(where, charge, n2p)
    character(len=*), intent(in) :: where
    real(c_double), intent(in) :: charge(8), n2p(8)
    if (differs(charge, spc0%charge, 8)) call arg_mismatch(where, 'charge')
    if (differs(n2p, spc0%n2p, 8)) call arg_mismatch(where, 'n2p')
  end subroutine check_charges

  subroutine check_one_rank(where, iprocs, jprocs, kprocs, myid)
    character(len=*), intent(in) :: where
    integer, intent(in) :: iprocs, jprocs, kprocs, myid
    if (iprocs /= 1 .or. jprocs /= 1 .or. kprocs /= 1 .or. myid /= 0) &
      call arg_mismatch(where, 'iprocs/jprocs/kprocs/myid (one domain, one rank)')
  end subroutine check_one_rank

  ! The arguments AdvanceVelocity, AdvanceVelocityImp and AdvanceVelocityExp
  ! share (AdvanceVelocity.f90:6-23; Exp :128-146; Imp :239-257); v0 is in their
  ! argument lists but none of them reads it.
  subroutine check_velocity(where, Lx_min, Ly_min, Lz_min, dxyz, B0, ve, eta, tB, fluid, gyration, &
                            max_per_proc, species, nx, ny, nz)
    character(len=*), intent(in) :: where
    real(c_double), intent(in) :: Lx_min, Ly_min, Lz_min, dxyz(3), B0(4), ve(4), eta(8), tB(32)
    logical, intent(in) :: fluid(8), gyration(8)
    integer, intent(in) :: max_per_proc, species, nx, ny, nz
    if (differs([Lx_min, Ly_min, Lz_min], geo0%lmin, 3)) call arg_mismatch(where, 'Lx_min/Ly_min/Lz_min')
    call check_dxyz(where, dxyz)
    call check_push(where, B0, ve, eta, tB, fluid, gyration)
    call check_capacity(where, max_per_proc)
    call check_species(where, species)
    call check_cells(where, nx, ny, nz)
  end subroutine check_velocity

  subroutine check_push(where, B0, ve, eta, tB, fluid, gyration)
    character(len=*), intent(in) :: where
    real(c_double), intent(in) :: B0(4), ve(4), eta(8), tB(32)
    logical, intent(in) :: fluid(8), gyration(8)
    integer :: s
    if (differs(B0, spc0%B0, 4)) call arg_mismatch(where, 'B0')
    if (differs(ve, spc0%ve, 4)) call arg_mismatch(where, 've')
    if (differs(eta, spc0%eta, 8)) call arg_mismatch(where, 'eta')
    if (differs(tB, spc0%tB, 32)) call arg_mismatch(where, 'tB')
    do s = 1, 8
      if (fluid(s) .neqv. (spc0%fluid(s) /= 0)) call arg_mismatch(where, 'fluid')
      if (gyration(s) .neqv. (spc0%gyration(s) /= 0)) call arg_mismatch(where, 'gyration')
    end do
  end subroutine check_push

  subroutine picard_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc < 0) then
      write(*,'(a,a,a,i0,a,a)') 'picard: ', where, ' failed (status ', rc, '): ', picard_error_message()
      error stop 1
    end if
  end subroutine picard_check

  ! Create the device domain and register the host particle arrays.
  subroutine picard_dropin_init(geo, spc, real_particles, int_particles, max_per_proc)
    type(picard_geometry), intent(in) :: geo
    type(picard_species), intent(in) :: spc
    integer, intent(in) :: max_per_proc
    real*8, intent(in), target :: real_particles(8, max_per_proc)
    integer, intent(in), target :: int_particles(max_per_proc)
    if (c_associated(ctx)) call picard_check(picard_destroy(ctx), 'picard_destroy')
    call picard_check(picard_create(geo, spc, ctx), 'picard_create')
    geo0 = geo
    spc0 = spc
    host_rp = c_loc(real_particles)
    host_sp = c_loc(int_particles)
    host_max = max_per_proc
    dev_count = 0
  end subroutine picard_dropin_init

  subroutine picard_dropin_finalize()
    if (c_associated(ctx)) call picard_check(picard_destroy(ctx), 'picard_destroy')
    ctx = c_null_ptr
  end subroutine picard_dropin_finalize

  ! Upload what the host appended after the device's count.
  subroutine picard_dropin_sync(num_local)
    integer, intent(in) :: num_local
    real(c_double), pointer :: rp(:,:)
    integer(c_int32_t), pointer :: sp(:)
    integer(c_int64_t) :: first, count
    if (.not. c_associated(ctx)) then
      write(*,*) 'picard: call picard_dropin_init before the time-step subroutines'
      error stop 1
    end if
    if (num_local > host_max) then
      write(*,*) 'picard: num_local exceeds max_per_proc'
      error stop 1
    end if
    if (num_local < dev_count) then
      write(*,*) 'picard: num_local shrank on the host; the device holds', dev_count
      error stop 1
    end if
    count = int(num_local, c_int64_t) - dev_count
    if (count == 0) return
    call c_f_pointer(host_rp, rp, [8, host_max])
    call c_f_pointer(host_sp, sp, [host_max])
    first = dev_count + 1
    call picard_check(picard_particles_append(ctx, c_loc(rp(1, first)), c_loc(sp(first)), count), &
                      'particles_append')
    dev_count = int(num_local, c_int64_t)
  end subroutine picard_dropin_sync

  ! Copy device particles back into the registered host arrays.
  subroutine picard_dropin_download_particles()
    real(c_double), pointer :: rp(:,:)
    integer(c_int32_t), pointer :: sp(:)
    if (dev_count == 0) return
    call c_f_pointer(host_rp, rp, [8, host_max])
    call c_f_pointer(host_sp, sp, [host_max])
    call picard_check(picard_particles_read(ctx, 0_c_int64_t, dev_count, c_loc(rp), c_loc(sp)), &
                      'particles_read')
  end subroutine picard_dropin_download_particles

  subroutine picard_dropin_download_field(which, a, n)
    integer(c_int32_t), intent(in) :: which
    integer(c_int64_t), intent(in) :: n
    real(c_double), intent(inout), target :: a(n)
    call picard_check(picard_field_read(ctx, which, c_loc(a), n), 'field_read')
  end subroutine picard_dropin_download_field

end module picard_dropin_state


! AdvancePosition.f90:4
subroutine AdvancePosition(real_particles, dt, max_per_proc, num_local)
  use picard_dropin_state
  implicit none
  integer, intent(in) :: max_per_proc, num_local
  real*8, intent(inout) :: real_particles(8, max_per_proc)
  real*8, intent(in) :: dt
  call check_capacity('AdvancePosition', max_per_proc)
  call picard_dropin_sync(num_local)
  call picard_check(picard_advance_position(ctx, dt), 'AdvancePosition')
  if (eager_download) call picard_dropin_download_particles()
end subroutine AdvancePosition


! BCoutflow.f90:5
subroutine BCoutflow(real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                     xmin, xmax, ymin, ymax, zmin, zmax, max_per_proc, num_local, xflow, yflow, zflow)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: xflow, yflow, zflow
  real*8, intent(in) :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, xmin, xmax, ymin, ymax, zmin, zmax
  integer, intent(in) :: max_per_proc
  real*8, intent(inout) :: real_particles(8, max_per_proc)
  integer, intent(inout) :: int_particles(max_per_proc), num_local
  integer(c_int64_t) :: n_new, removed
  call check_capacity('BCoutflow', max_per_proc)
  call check_local_box('BCoutflow', [Lx_min, Ly_min, Lz_min], [Lx_max, Ly_max, Lz_max])
  call check_global_box('BCoutflow', [xmin, ymin, zmin], [xmax, ymax, zmax])
  call check_flows('BCoutflow', xflow, yflow, zflow)
  call picard_dropin_sync(num_local)
  call picard_check(picard_bc_outflow(ctx, n_new, removed), 'BCoutflow')
  num_local = int(n_new)
  dev_count = n_new
  last_removed = removed
  if (eager_download) call picard_dropin_download_particles()
end subroutine BCoutflow


! ParticlesTransfer.f90:2 (one domain)
subroutine ParticlesTransfer(real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                             xmin, xmax, ymin, ymax, zmin, zmax, max_per_proc, num_local, iprocs, jprocs, &
                             kprocs, myid)
  use picard_dropin_state
  implicit none
  integer, intent(in) :: max_per_proc, iprocs, jprocs, kprocs, myid
  real*8, intent(in) :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, xmin, xmax, ymin, ymax, zmin, zmax
  integer, intent(inout) :: int_particles(max_per_proc), num_local
  real*8, intent(inout) :: real_particles(8, max_per_proc)
  if (iprocs /= 1 .or. jprocs /= 1 .or. kprocs /= 1 .or. myid /= 0) then
    write(*,*) 'picard: one domain per GPU, one rank (iprocs = jprocs = kprocs = 1, myid = 0)'
    error stop 1
  end if
  call check_capacity('ParticlesTransfer', max_per_proc)
  call check_local_box('ParticlesTransfer', [Lx_min, Ly_min, Lz_min], [Lx_max, Ly_max, Lz_max])
  call check_global_box('ParticlesTransfer', [xmin, ymin, zmin], [xmax, ymax, zmax])
  call picard_dropin_sync(num_local)
  call picard_check(picard_particles_transfer(ctx), 'ParticlesTransfer')
  if (eager_download) call picard_dropin_download_particles()
end subroutine ParticlesTransfer


! ParticlesGridify.f90:1
subroutine ParticlesGridify(N, C, real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, &
                            Lz_max, dxyz, xmin, xmax, ymin, ymax, zmin, zmax, species, Nx_local, Ny_local, &
                            Nz_local, max_per_proc, num_local, xflow, yflow, zflow, iprocs, jprocs, kprocs, &
                            charge, n2p, myid)
  use picard_dropin_state
  implicit none
  integer, intent(in) :: species, Nx_local, Ny_local, Nz_local, max_per_proc, iprocs, jprocs, kprocs, myid, &
                         num_local
  logical, intent(in) :: xflow, yflow, zflow
  real*8, intent(in) :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, dxyz(3), xmin, xmax, ymin, ymax, &
                        zmin, zmax, charge(8), n2p(8)
  integer, intent(in) :: int_particles(max_per_proc)
  real*8, intent(in) :: real_particles(8, max_per_proc)
  real*8, intent(out) :: N(species, Nx_local+2, Ny_local+2, Nz_local+2)
  real*8, intent(out) :: C(Nx_local+2, Ny_local+2, Nz_local+2)
  integer(c_int64_t) :: G
  call check_gridify('ParticlesGridify')
  call picard_dropin_sync(num_local)
  call picard_check(picard_particles_gridify(ctx), 'ParticlesGridify')
  if (eager_download) then
    G = int(Nx_local+2, c_int64_t) * (Ny_local+2) * (Nz_local+2)
    call picard_dropin_download_field(PICARD_FIELD_N, N, species * G)
    call picard_dropin_download_field(PICARD_FIELD_C, C, G)
  end if
contains
  subroutine check_gridify(where)
    character(len=*), intent(in) :: where
    call check_one_rank(where, iprocs, jprocs, kprocs, myid)
    call check_capacity(where, max_per_proc)
    call check_cells(where, Nx_local, Ny_local, Nz_local)
    call check_species(where, species)
    call check_local_box(where, [Lx_min, Ly_min, Lz_min], [Lx_max, Ly_max, Lz_max])
    call check_global_box(where, [xmin, ymin, zmin], [xmax, ymax, zmax])
    call check_dxyz(where, dxyz)
    call check_flows(where, xflow, yflow, zflow)
    call check_charges(where, charge, n2p)
  end subroutine check_gridify
end subroutine ParticlesGridify


! CalcDistance.f90:1 (global Nx = Nx_local on one domain)
subroutine CalcDistance(D, Nx, Ny, Nz, dxyz, xflow, yflow, zflow)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: xflow, yflow, zflow
  integer, intent(in) :: Nx, Ny, Nz
  real*8, intent(in) :: dxyz(3)
  real*8, intent(out) :: D(2*Nx+1, 2*Ny+1, 2*Nz+1)
  call check_cells('CalcDistance', Nx, Ny, Nz)
  call check_dxyz('CalcDistance', dxyz)
  call check_flows('CalcDistance', xflow, yflow, zflow)
  call picard_check(picard_calc_distance(ctx), 'CalcDistance')
  if (eager_download) &
    call picard_dropin_download_field(PICARD_FIELD_D, D, int(2*Nx+1, c_int64_t) * (2*Ny+1) * (2*Nz+1))
end subroutine CalcDistance


! BCpotentials.f90:1
subroutine BCpotentials(U, C, D, xmin, xmax, ymin, ymax, zmin, zmax, Nx_local, Ny_local, Nz_local, Nx, Ny, Nz, &
                        xflow, yflow, zflow, species, iprocs, jprocs, kprocs, myid)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: xflow, yflow, zflow
  integer, intent(in) :: Nx_local, Ny_local, Nz_local, Nx, Ny, Nz, species, iprocs, jprocs, kprocs, myid
  real*8, intent(in) :: xmin, xmax, ymin, ymax, zmin, zmax
  real*8, intent(in) :: D(2*Nx+1, 2*Ny+1, 2*Nz+1)
  real*8, intent(in) :: C(Nx_local+2, Ny_local+2, Nz_local+2)
  real*8, intent(inout) :: U(Nx_local+2, Ny_local+2, Nz_local+2)
  call check_one_rank('BCpotentials', iprocs, jprocs, kprocs, myid)
  call check_cells('BCpotentials', Nx_local, Ny_local, Nz_local)
  call check_cells('BCpotentials', Nx, Ny, Nz)
  call check_global_box('BCpotentials', [xmin, ymin, zmin], [xmax, ymax, zmax])
  call check_flows('BCpotentials', xflow, yflow, zflow)
  call check_species('BCpotentials', species)
  call picard_check(picard_bc_potentials(ctx), 'BCpotentials')
  if (eager_download) &
    call picard_dropin_download_field(PICARD_FIELD_U, U, int(Nx_local+2, c_int64_t) * (Ny_local+2) * (Nz_local+2))
end subroutine BCpotentials


! CalcEpotential.f90:4
subroutine CalcEpotential(U, C, Nx_local, Ny_local, Nz_local, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, &
                          xflow, yflow, zflow, xmin, xmax, ymin, ymax, zmin, zmax, dxyz, maxerr, maxit, &
                          iprocs, jprocs, kprocs, species, myid)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: xflow, yflow, zflow
  integer, intent(in) :: Nx_local, Ny_local, Nz_local, maxit, iprocs, jprocs, kprocs, species, myid
  real*8, intent(in) :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, xmin, xmax, ymin, ymax, zmin, zmax, &
                        maxerr, dxyz(3)
  real*8, intent(in) :: C(Nx_local+2, Ny_local+2, Nz_local+2)
  real*8, intent(inout) :: U(Nx_local+2, Ny_local+2, Nz_local+2)
  integer(c_int) :: rc
  call check_one_rank('CalcEpotential', iprocs, jprocs, kprocs, myid)
  call check_cells('CalcEpotential', Nx_local, Ny_local, Nz_local)
  call check_local_box('CalcEpotential', [Lx_min, Ly_min, Lz_min], [Lx_max, Ly_max, Lz_max])
  call check_global_box('CalcEpotential', [xmin, ymin, zmin], [xmax, ymax, zmax])
  call check_flows('CalcEpotential', xflow, yflow, zflow)
  call check_dxyz('CalcEpotential', dxyz)
  call check_species('CalcEpotential', species)
  rc = picard_calc_epotential(ctx, maxerr, int(maxit, c_int32_t), last_sweeps)
  call picard_check(rc, 'CalcEpotential')
  if (rc == PICARD_WARN_NOCONV) write(*,*) 'Maximum iteration reached!'
  if (eager_download) &
    call picard_dropin_download_field(PICARD_FIELD_U, U, int(Nx_local+2, c_int64_t) * (Ny_local+2) * (Nz_local+2))
end subroutine CalcEpotential


! CalcEfield.f90:3
subroutine CalcEfield(E, U, Nx_local, Ny_local, Nz_local, dxyz, xflow, yflow, zflow, iprocs, jprocs, kprocs, myid)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: xflow, yflow, zflow
  integer, intent(in) :: Nx_local, Ny_local, Nz_local, iprocs, jprocs, kprocs, myid
  real*8, intent(in) :: U(Nx_local+2, Ny_local+2, Nz_local+2), dxyz(3)
  real*8, intent(out) :: E(3, Nx_local+2, Ny_local+2, Nz_local+2)
  call check_one_rank('CalcEfield', iprocs, jprocs, kprocs, myid)
  call check_cells('CalcEfield', Nx_local, Ny_local, Nz_local)
  call check_dxyz('CalcEfield', dxyz)
  call check_flows('CalcEfield', xflow, yflow, zflow)
  call picard_check(picard_calc_efield(ctx), 'CalcEfield')
  if (eager_download) &
    call picard_dropin_download_field(PICARD_FIELD_E, E, 3_c_int64_t * (Nx_local+2) * (Ny_local+2) * (Nz_local+2))
end subroutine CalcEfield


! BCfields.f90:3
subroutine BCfields(E, U, B0, E0, Nx_local, Ny_local, Nz_local, xflow, yflow, zflow, iprocs, jprocs, kprocs, myid)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: xflow, yflow, zflow
  integer, intent(in) :: Nx_local, Ny_local, Nz_local, iprocs, jprocs, kprocs, myid
  real*8 :: B0(4), E0(4)
  real*8, intent(in) :: U(Nx_local+2, Ny_local+2, Nz_local+2)
  real*8, intent(inout) :: E(3, Nx_local+2, Ny_local+2, Nz_local+2)
  call check_one_rank('BCfields', iprocs, jprocs, kprocs, myid)
  call check_cells('BCfields', Nx_local, Ny_local, Nz_local)
  call check_flows('BCfields', xflow, yflow, zflow)
  if (differs(B0, spc0%B0, 4)) call arg_mismatch('BCfields', 'B0')
  if (differs(E0, spc0%E0, 4)) call arg_mismatch('BCfields', 'E0')
  call picard_check(picard_bc_fields(ctx), 'BCfields')
  if (eager_download) &
    call picard_dropin_download_field(PICARD_FIELD_E, E, 3_c_int64_t * (Nx_local+2) * (Ny_local+2) * (Nz_local+2))
end subroutine BCfields


! AdvanceVelocity.f90:6
subroutine AdvanceVelocity(real_particles, int_particles, E, B0, Lx_min, Ly_min, Lz_min, dxyz, dt, v0, ve, eta, &
                           tB, sB, num_local, max_per_proc, species, Nx_local, Ny_local, Nz_local, fluid, gyration)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: fluid(8), gyration(8)
  integer, intent(in) :: num_local, max_per_proc, species, Nx_local, Ny_local, Nz_local
  real*8, intent(in) :: Lx_min, Ly_min, Lz_min, dxyz(3), dt, v0(4,8), ve(4), eta(8)
  real*8, intent(in) :: tB(4,8), sB(4,8)
  real*8, intent(inout) :: real_particles(8, max_per_proc)
  integer, intent(in) :: int_particles(max_per_proc)
  real*8, intent(in) :: B0(4)
  real*8, intent(in) :: E(3, Nx_local+2, Ny_local+2, Nz_local+2)
  call check_velocity('AdvanceVelocity', Lx_min, Ly_min, Lz_min, dxyz, B0, ve, eta, tB, fluid, gyration, &
                      max_per_proc, species, Nx_local, Ny_local, Nz_local)
  if (differs(sB, spc0%sB, 32)) call arg_mismatch('AdvanceVelocity', 'sB')
  call picard_dropin_sync(num_local)
  call picard_check(picard_advance_velocity(ctx, dt), 'AdvanceVelocity')
  if (eager_download) call picard_dropin_download_particles()
end subroutine AdvanceVelocity


! AdvanceVelocity.f90:239 (output steps, before the dumps)
subroutine AdvanceVelocityImp(real_particles, int_particles, E, B0, Lx_min, Ly_min, Lz_min, dxyz, deltat, dt, &
                              v0, ve, eta, tB, num_local, max_per_proc, species, Nx_local, Ny_local, Nz_local, &
                              fluid, gyration)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: fluid(8), gyration(8)
  integer, intent(in) :: num_local, max_per_proc, species, Nx_local, Ny_local, Nz_local
  real*8, intent(in) :: Lx_min, Ly_min, Lz_min, dxyz(3), deltat, dt, v0(4,8), ve(4), eta(8)
  real*8, intent(in) :: tB(4,8)
  real*8, intent(inout) :: real_particles(8, max_per_proc)
  integer, intent(in) :: int_particles(max_per_proc)
  real*8, intent(in) :: B0(4)
  real*8, intent(in) :: E(3, Nx_local+2, Ny_local+2, Nz_local+2)
  call check_velocity('AdvanceVelocityImp', Lx_min, Ly_min, Lz_min, dxyz, B0, ve, eta, tB, fluid, gyration, &
                      max_per_proc, species, Nx_local, Ny_local, Nz_local)
  call picard_dropin_sync(num_local)
  call picard_check(picard_advance_velocity_imp(ctx, deltat, dt), 'AdvanceVelocityImp')
  if (eager_download) call picard_dropin_download_particles()
end subroutine AdvanceVelocityImp


! AdvanceVelocity.f90:128 (output steps, after the dumps)
subroutine AdvanceVelocityExp(real_particles, int_particles, E, B0, Lx_min, Ly_min, Lz_min, dxyz, deltat, dt, &
                              v0, ve, eta, tB, num_local, max_per_proc, species, Nx_local, Ny_local, Nz_local, &
                              fluid, gyration)
  use picard_dropin_state
  implicit none
  logical, intent(in) :: fluid(8), gyration(8)
  integer, intent(in) :: num_local, max_per_proc, species, Nx_local, Ny_local, Nz_local
  real*8, intent(in) :: Lx_min, Ly_min, Lz_min, dxyz(3), deltat, dt, v0(4,8), ve(4), eta(8)
  real*8, intent(in) :: tB(4,8)
  real*8, intent(inout) :: real_particles(8, max_per_proc)
  integer, intent(in) :: int_particles(max_per_proc)
  real*8, intent(in) :: B0(4)
  real*8, intent(in) :: E(3, Nx_local+2, Ny_local+2, Nz_local+2)
  call check_velocity('AdvanceVelocityExp', Lx_min, Ly_min, Lz_min, dxyz, B0, ve, eta, tB, fluid, gyration, &
                      max_per_proc, species, Nx_local, Ny_local, Nz_local)
  call picard_dropin_sync(num_local)
  call picard_check(picard_advance_velocity_exp(ctx, deltat, dt), 'AdvanceVelocityExp')
  if (eager_download) call picard_dropin_download_particles()
end subroutine AdvanceVelocityExp


! ParticlesGridifyAll.f90:1 (field-output steps).  The dumps read F, N and C
! on the host, so they are always copied back.
subroutine ParticlesGridifyAll(F, N, C, real_particles, int_particles, Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, &
                               Lz_max, dxyz, xmin, xmax, ymin, ymax, zmin, zmax, species, Nx_local, Ny_local, &
                               Nz_local, max_per_proc, num_local, xflow, yflow, zflow, iprocs, jprocs, kprocs, &
                               charge, n2p, myid)
  use picard_dropin_state
  implicit none
  integer, intent(in) :: species, Nx_local, Ny_local, Nz_local, max_per_proc, iprocs, jprocs, kprocs, myid, &
                         num_local
  logical, intent(in) :: xflow, yflow, zflow
  real*8, intent(in) :: Lx_min, Lx_max, Ly_min, Ly_max, Lz_min, Lz_max, dxyz(3), xmin, xmax, ymin, ymax, &
                        zmin, zmax, charge(8), n2p(8)
  integer, intent(in) :: int_particles(max_per_proc)
  real*8, intent(in) :: real_particles(8, max_per_proc)
  real*8, intent(out) :: F(species, 3, Nx_local+2, Ny_local+2, Nz_local+2)
  real*8, intent(out) :: N(species, Nx_local+2, Ny_local+2, Nz_local+2)
  real*8, intent(out) :: C(Nx_local+2, Ny_local+2, Nz_local+2)
  integer(c_int64_t) :: G
  call check_one_rank('ParticlesGridifyAll', iprocs, jprocs, kprocs, myid)
  call check_capacity('ParticlesGridifyAll', max_per_proc)
  call check_cells('ParticlesGridifyAll', Nx_local, Ny_local, Nz_local)
  call check_species('ParticlesGridifyAll', species)
  call check_local_box('ParticlesGridifyAll', [Lx_min, Ly_min, Lz_min], [Lx_max, Ly_max, Lz_max])
  call check_global_box('ParticlesGridifyAll', [xmin, ymin, zmin], [xmax, ymax, zmax])
  call check_dxyz('ParticlesGridifyAll', dxyz)
  call check_flows('ParticlesGridifyAll', xflow, yflow, zflow)
  call check_charges('ParticlesGridifyAll', charge, n2p)
  call picard_dropin_sync(num_local)
  call picard_check(picard_particles_gridify_all(ctx), 'ParticlesGridifyAll')
  G = int(Nx_local+2, c_int64_t) * (Ny_local+2) * (Nz_local+2)
  call picard_dropin_download_field(PICARD_FIELD_F, F, 3 * species * G)
  call picard_dropin_download_field(PICARD_FIELD_N, N, species * G)
  call picard_dropin_download_field(PICARD_FIELD_C, C, G)
end subroutine ParticlesGridifyAll


! The diagnostics line of the main loop (picard.f90:649-672) from device
! state: num_global and the particle, electric and magnetic energies.
subroutine picard_dropin_diagnostics(mass, num_global, uP, uE, uB)
  use picard_dropin_state
  implicit none
  real*8, intent(in) :: mass(8)
  integer, intent(out) :: num_global
  real*8, intent(out) :: uP, uE, uB
  type(picard_diag) :: dg
  call picard_check(picard_diagnostics(ctx, mass, dg), 'diagnostics')
  num_global = int(dg%num_global)
  uP = dg%particle_energy
  uE = dg%e_energy
  uB = dg%b_energy
end subroutine picard_dropin_diagnostics
