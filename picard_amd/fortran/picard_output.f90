! picard_output.f90 -- text dumps of particles and fields in the reference's
! on-disk layout (DumpParticles and DumpFields, Write.f90:2-523): edit
! descriptor E10.4, particles seven values per record (positions, velocities,
! species), fields one value per record in array order with the guard layers,
! files save/particles/particles_IIIII_RRRRR and save/fields/<name>_IIIII_RRRRR
! (name UE, Ex, Ey, Ez, N<s>, X<s>, Y<s>, Z<s> for s = 1..species).
module picard_output
  implicit none
  private
  public :: dump_particles, dump_fields

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

end module picard_output
