! picard_output.f90 -- text dumps of particles and fields in the reference's
! on-disk layout (DumpParticles and DumpFields, Write.f90:2-523): edit
! descriptor E10.4, particles seven values per record (positions, velocities,
! species), fields one value per record in array order with the guard layers,
! files save/particles/particles_IIIII_RRRRR and save/fields/<name>_IIIII_RRRRR
! (name UE, Ex, Ey, Ez, N<s>, X<s>, Y<s>, Z<s> for s = 1..species); and the
! run's save/parameters.dat (WriteWarnings.f90:130-256).
module picard_output
  implicit none
  private
  public :: dump_particles, dump_fields, dump_parameters

contains

  subroutine open_dump(u, dir, kind, iteration, rank)
    integer, intent(out) :: u
    character(len=*), intent(in) :: dir, kind
    integer, intent(in) :: iteration, rank
    character(len=5) :: it, rk
    write(it, '(I5.5)') iteration
    write(rk, '(I5.5)') rank
    open(newunit=u, file=trim(dir)//'/'//trim(kind)//'_'//it//'_'//rk, status='replace', action='write')
  end subroutine open_dump

  ! Particles 1..num_local; slot 7 of the record is the species.
  subroutine dump_particles(real_particles, int_particles, num_local, iteration, rank)
    real*8, intent(in) :: real_particles(:,:)
    integer, intent(in) :: int_particles(:)
    integer, intent(in) :: num_local, iteration, rank
    integer :: u, p
    real*8 :: rec(7)
    call execute_command_line('mkdir -p save/particles')
    call open_dump(u, 'save/particles', 'particles', iteration, rank)
    do p = 1, num_local
      rec(1:6) = real_particles(1:6, p)
      rec(7) = dble(int_particles(p))
      write(u, fmt='(7E10.4)') rec
    end do
    close(u)
  end subroutine dump_particles

  subroutine dump_one(dir, kind, a, iteration, rank)
    character(len=*), intent(in) :: dir, kind
    real*8, intent(in) :: a(:,:,:)
    integer, intent(in) :: iteration, rank
    integer :: u, i, j, k
    call open_dump(u, dir, kind, iteration, rank)
    do k = 1, size(a, 3)
      do j = 1, size(a, 2)
        do i = 1, size(a, 1)
          write(u, fmt='(6E10.4)') a(i, j, k)
        end do
      end do
    end do
    close(u)
  end subroutine dump_one

  subroutine dump_fields(U, N, F, E, species, iteration, rank)
    real*8, intent(in) :: U(:,:,:), N(:,:,:,:), F(:,:,:,:,:), E(:,:,:,:)
    integer, intent(in) :: species, iteration, rank
    character(len=1), parameter :: comp(3) = ['X', 'Y', 'Z']
    character(len=2) :: nm
    integer :: s, c
    call execute_command_line('mkdir -p save/fields')
    call dump_one('save/fields', 'UE', U, iteration, rank)
    call dump_one('save/fields', 'Ex', E(1,:,:,:), iteration, rank)
    call dump_one('save/fields', 'Ey', E(2,:,:,:), iteration, rank)
    call dump_one('save/fields', 'Ez', E(3,:,:,:), iteration, rank)
    do s = 1, species
      write(nm, '(A1,I1)') 'N', s
      call dump_one('save/fields', nm, N(s,:,:,:), iteration, rank)
    end do
    do c = 1, 3
      do s = 1, species
        write(nm, '(A1,I1)') comp(c), s
        call dump_one('save/fields', nm, F(s,c,:,:,:), iteration, rank)
      end do
    end do
  end subroutine dump_fields

  ! save/parameters.dat as the reference writes it at setup
  ! (WriteWarnings.f90:130-256, picard.f90:488-489): one `name = value`
  ! record per quantity in the reference's order, reals in E12.6, integers in
  ! I4, logicals list-directed.  bounds = xmin, xmax, ymin, ymax, zmin, zmax;
  ! ve is the bulk velocity v0 of the file.
  subroutine dump_parameters(n2p, mass, charge, density, kelvin, Qn, vn, Ek, nu_i, nu_d, ppc, &
                             gyration, ve, B0, dxyz, bounds, dt, nxyz, procs, flow)
    real*8, intent(in) :: n2p(8), mass(8), charge(8), density(8), kelvin(8), Qn(8), vn(8), &
                          Ek(8), nu_i(8), nu_d(8), ve(4), B0(4), dxyz(3), bounds(6), dt
    integer, intent(in) :: ppc(8), nxyz(3), procs(3)
    logical, intent(in) :: gyration(8), flow(3)
    character(len=6), parameter :: bname(6) = [character(len=6) :: 'xmin', 'xmax', 'ymin', &
                                                'ymax', 'zmin', 'zmax']
    character(len=6), parameter :: iname(6) = [character(len=6) :: 'Nx', 'Ny', 'Nz', 'iprocs', &
                                                'jprocs', 'kprocs']
    character(len=1), parameter :: axis(3) = ['x', 'y', 'z']
    integer :: u, s, q(6)
    call execute_command_line('mkdir -p save')
    open(newunit=u, file='save/parameters.dat', status='replace', action='write')
    call per_species(u, 'n2p', n2p)
    call per_species(u, 'mass', mass)
    call per_species(u, 'charge', charge)
    call per_species(u, 'density', density)
    call per_species(u, 'kelvin', kelvin)
    call per_species(u, 'Qn', Qn)
    call per_species(u, 'vn', vn)
    call per_species(u, 'Ek', Ek)
    call per_species(u, 'nui', nu_i)
    call per_species(u, 'nud', nu_d)
    do s = 1, 8
      write(u, '(a,i1,a,I4)') 'ppc_', s, ' = ', ppc(s)
    end do
    do s = 1, 8
      write(u, *) 'gyration_'//achar(48 + s)//' = ', gyration(s)
    end do
    do s = 1, 3
      write(u, '(a,i1,a,E12.6)') 'v0_', s, ' = ', ve(s)
    end do
    do s = 1, 3
      write(u, '(a,i1,a,E12.6)') 'B0_', s, ' = ', B0(s)
    end do
    do s = 1, 3
      write(u, '(a,i1,a,E12.6)') 'dxyz_', s, ' = ', dxyz(s)
    end do
    do s = 1, 6
      write(u, '(a,a,E12.6)') trim(bname(s)), ' = ', bounds(s)
    end do
    write(u, '(a,E12.6)') 'dt = ', dt
    q(1:3) = nxyz
    q(4:6) = procs
    do s = 1, 6
      write(u, '(a,a,I4)') trim(iname(s)), ' = ', q(s)
    end do
    do s = 1, 3
      write(u, *) axis(s)//'flow = ', flow(s)
    end do
    close(u)
  contains
    subroutine per_species(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      real*8, intent(in) :: a(8)
      integer :: k
      do k = 1, 8
        write(u, '(a,a,i1,a,E12.6)') name, '_', k, ' = ', a(k)
      end do
    end subroutine per_species
  end subroutine dump_parameters

end module picard_output
