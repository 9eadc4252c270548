! picard_hip_iface.f90 -- ISO_C_BINDING interfaces to libpicard_hip.so
! (include/picard_hip.h).  The derived types mirror the C structs field for
! field; every function returns the C status code.
module picard_hip_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PICARD_OK = 0
  integer(c_int), parameter :: PICARD_WARN_NOCONV = 1
  integer(c_int32_t), parameter :: PICARD_FP64 = 0
  integer(c_int32_t), parameter :: PICARD_DEPOSIT_REFERENCE_ORDER = 0
  integer(c_int32_t), parameter :: PICARD_FIELD_U = 0, PICARD_FIELD_E = 1, &
                                   PICARD_FIELD_N = 2, PICARD_FIELD_C = 3, PICARD_FIELD_D = 4, &
                                   PICARD_FIELD_F = 5

  type, bind(C) :: picard_geometry
    integer(c_int32_t) :: nx, ny, nz, species
    integer(c_int32_t) :: flow(3)
    integer(c_int32_t) :: precision, deposit, device
    integer(c_int64_t) :: max_particles
    real(c_double) :: lmin(3), lmax(3), gmin(3), gmax(3), dxyz(3)
  end type picard_geometry

  type, bind(C) :: picard_species
    real(c_double) :: charge(8), n2p(8), eta(8)
    real(c_double) :: tB(32), sB(32)
    real(c_double) :: ve(4), B0(4), E0(4)
    integer(c_int32_t) :: fluid(8), gyration(8)
  end type picard_species

  type, bind(C) :: picard_diag
    integer(c_int64_t) :: num_local, num_global
    real(c_double) :: particle_energy, e_energy, b_energy
  end type picard_diag

  interface
    function picard_last_error() bind(C, name='picard_last_error')
      import :: c_ptr
      type(c_ptr) :: picard_last_error
    end function

    integer(c_int) function picard_create(geo, sp, ctx) bind(C, name='picard_create')
      import :: c_int, c_ptr, picard_geometry, picard_species
      type(picard_geometry), intent(in) :: geo
      type(picard_species), intent(in) :: sp
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function picard_destroy(ctx) bind(C, name='picard_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_particles_append(ctx, rp, sp, count) &
        bind(C, name='picard_particles_append')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, rp, sp
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function picard_particles_read(ctx, first, count, rp, sp) &
        bind(C, name='picard_particles_read')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first, count
      type(c_ptr), value :: rp, sp
    end function

    integer(c_int) function picard_advance_position(ctx, dt) bind(C, name='picard_advance_position')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function

    integer(c_int) function picard_bc_outflow(ctx, num_local, removed) bind(C, name='picard_bc_outflow')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: num_local, removed
    end function

    integer(c_int) function picard_particles_transfer(ctx) bind(C, name='picard_particles_transfer')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_particles_gridify(ctx) bind(C, name='picard_particles_gridify')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_calc_distance(ctx) bind(C, name='picard_calc_distance')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_bc_potentials(ctx) bind(C, name='picard_bc_potentials')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_calc_epotential(ctx, maxerr, maxit, sweeps) &
        bind(C, name='picard_calc_epotential')
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: maxerr
      integer(c_int32_t), value :: maxit
      integer(c_int64_t), intent(out) :: sweeps
    end function

    integer(c_int) function picard_calc_efield(ctx) bind(C, name='picard_calc_efield')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_bc_fields(ctx) bind(C, name='picard_bc_fields')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_advance_velocity(ctx, dt) bind(C, name='picard_advance_velocity')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function

    integer(c_int) function picard_advance_velocity_imp(ctx, deltat, dt) &
        bind(C, name='picard_advance_velocity_imp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, dt
    end function

    integer(c_int) function picard_advance_velocity_exp(ctx, deltat, dt) &
        bind(C, name='picard_advance_velocity_exp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, dt
    end function

    integer(c_int) function picard_particles_gridify_all(ctx) bind(C, name='picard_particles_gridify_all')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function picard_diagnostics(ctx, mass, out) bind(C, name='picard_diagnostics')
      import :: c_int, c_ptr, c_double, picard_diag
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: mass(8)
      type(picard_diag), intent(out) :: out
    end function

    integer(c_int) function picard_field_read(ctx, which, host, count) bind(C, name='picard_field_read')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      type(c_ptr), value :: host
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function picard_field_write(ctx, which, host, count) bind(C, name='picard_field_write')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      type(c_ptr), value :: host
      integer(c_int64_t), value :: count
    end function
  end interface

contains

  function picard_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
      end function
    end interface
    p = picard_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function picard_error_message

end module picard_hip_iface
