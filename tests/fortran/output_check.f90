! Writes the Fortran host's dumps (module picard_output) from raw inputs, for
! tests/test_output.py to compare with the reference's files.
!   output_check <dir>          reads <dir>/dims.txt and <dir>/*.bin (stream,
!                               FP64), writes <dir>/save/...
!   output_check <dir> params   reads <dir>/params.bin, writes
!                               <dir>/save/parameters.dat
program output_check
  use picard_output
  implicit none
  character(len=512) :: dir
  integer :: n, nx, ny, nz, s, u, it_p, it_f
  real*8, allocatable :: rp(:,:), U3(:,:,:), E(:,:,:,:), N4(:,:,:,:), F(:,:,:,:,:)
  integer, allocatable :: sp(:)
  character(len=16) :: mode
  call get_command_argument(1, dir)
  call get_command_argument(2, mode)
  if (trim(mode) == 'params') then
    call params(trim(dir))
    stop
  end if
  open(newunit=u, file=trim(dir)//'/dims.txt', action='read')
  read(u, *) n, nx, ny, nz, s, it_p, it_f
  close(u)
  allocate(rp(8, n), sp(n), U3(nx+2, ny+2, nz+2), E(3, nx+2, ny+2, nz+2))
  allocate(N4(s, nx+2, ny+2, nz+2), F(s, 3, nx+2, ny+2, nz+2))
  call rd(trim(dir)//'/rp.bin', rp, size(rp))
  open(newunit=u, file=trim(dir)//'/sp.bin', access='stream', form='unformatted', action='read')
  read(u) sp
  close(u)
  call rd(trim(dir)//'/U.bin', U3, size(U3))
  call rd(trim(dir)//'/E.bin', E, size(E))
  call rd(trim(dir)//'/N.bin', N4, size(N4))
  call rd(trim(dir)//'/F.bin', F, size(F))
  call chdir_to(trim(dir))
  call dump_particles(rp, sp, n, it_p, 0)
  call dump_fields(U3, N4, F, E, s, it_f, 0)
contains
  ! params.bin: 10 x 8 FP64 (n2p, mass, charge, density, kelvin, Qn, vn, Ek,
  ! nu_i, nu_d), ve(4), B0(4), dxyz(3), bounds(6), dt as FP64, then int32
  ! ppc(8), gyration(8), nxyz(3), procs(3), flow(3) (logicals as 0/1)
  subroutine params(path)
    character(len=*), intent(in) :: path
    real*8 :: a(8, 10), ve(4), B0(4), dxyz(3), bounds(6), dt
    integer :: ppc(8), gy(8), nxyz(3), procs(3), fl(3), v
    open(newunit=v, file=path//'/params.bin', access='stream', form='unformatted', action='read')
    read(v) a, ve, B0, dxyz, bounds, dt, ppc, gy, nxyz, procs, fl
    close(v)
    call chdir_to(path)
    call dump_parameters(a(:,1), a(:,2), a(:,3), a(:,4), a(:,5), a(:,6), a(:,7), a(:,8), &
                         a(:,9), a(:,10), ppc, gy /= 0, ve, B0, dxyz, bounds, dt, nxyz, procs, fl /= 0)
  end subroutine params
  subroutine rd(path, a, m)
    character(len=*), intent(in) :: path
    integer, intent(in) :: m
    real*8, intent(out) :: a(m)
    integer :: v
    open(newunit=v, file=path, access='stream', form='unformatted', action='read')
    read(v) a
    close(v)
  end subroutine rd
  subroutine chdir_to(path)
    use iso_c_binding
    character(len=*), intent(in) :: path
    interface
      integer(c_int) function c_chdir(p) bind(C, name='chdir')
        import :: c_int, c_char
        character(kind=c_char), intent(in) :: p(*)
      end function
    end interface
    if (c_chdir(path//c_null_char) /= 0) error stop 1
  end subroutine chdir_to
end program output_check
