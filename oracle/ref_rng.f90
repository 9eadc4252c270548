! Seeding and drawing from the Fortran runtime's random_number, the generator
! the reference's host sources use (ParticlesProduce.f90:69, :74).  Built only
! into oracle/_ref beside the reference's own subroutines, so a test can
! replay the exact uniform stream a reference call consumed.
subroutine picard_ref_seed(seed) bind(C, name="picard_ref_seed")
  use iso_c_binding
  implicit none
  integer(c_int32_t), value :: seed
  integer :: k
  integer, allocatable :: s(:)
  call random_seed(size=k)
  allocate(s(k))
  s = seed
  call random_seed(put=s)
  deallocate(s)
end subroutine picard_ref_seed

subroutine picard_ref_uniforms(out, n) bind(C, name="picard_ref_uniforms")
  use iso_c_binding
  implicit none
  integer(c_int64_t), value :: n
  real(c_double), intent(out) :: out(n)
  call random_number(out)
end subroutine picard_ref_uniforms

! Flush the runtime's standard output (the reference's list-directed
! summaries), so a caller that redirects file descriptor 1 receives them.
subroutine picard_ref_flush() bind(C, name='picard_ref_flush')
  flush(6)
end subroutine picard_ref_flush
