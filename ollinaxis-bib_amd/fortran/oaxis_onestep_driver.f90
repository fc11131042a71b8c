! Fortran main loop that keeps the reference's onestep recursion
! (onestep.f90:78-469) and replaces only its body with the piecewise entry
! points of liboaxis.so: oa_saveold, then per internal iteration oa_substep,
! oa_finebound (fine levels) and oa_auxiliary, then oa_advance_time, the
! recursion into the next level twice, and restriction with the parent's
! auxiliary when the fine level's step count is even.
! Usage:  oaxis_onestep_driver <parfile> <coarse steps> <output file> <array names...>
! Writes every named array of every level, level by level and in the order
! given, as raw doubles in the reference layout (1-ghost:Nr, 1-ghost:Nz),
! column-major, and prints the level clocks.
module onestep_body
  use, intrinsic :: iso_c_binding
  use oaxis_mod
  implicit none
  type(c_ptr) :: h
  integer(c_int) :: nlev, niter
  logical :: rk4
contains

  recursive subroutine onestep(level)
    integer(c_int), intent(in) :: level
    integer(c_int) :: k, Nr, Nz, ghost, s
    real(c_double) :: dr, dz, dt, dtw, w, t, t1, t2
    call oaxis_check(oa_level_info(h, 0_c_int, level, Nr, Nz, ghost, dr, dz, dt), "oa_level_info")
    call oaxis_check(oa_saveold(h, 0_c_int, level), "oa_saveold")                  ! :97
    do k = 1, niter                                                                ! :129-342
      ! dtw and the weight as onestep.f90:141-177 forms them
      w = 0.0d0
      if (rk4) then
        select case (k)
        case (1)
          dtw = 0.5d0*dt; w = 1.0d0/6.0d0
        case (2)
          dtw = 0.5d0*dt; w = 1.0d0/3.0d0
        case (3)
          dtw = dt; w = 1.0d0/3.0d0
        case default
          dtw = dt; w = 1.0d0/6.0d0
        end select
      else
        if (k < niter) then
          dtw = 0.5d0*dt
        else
          dtw = dt
        end if
      end if
      call oaxis_check(oa_substep(h, 0_c_int, level, k, niter, dtw, w), "oa_substep")
      if (level > 0) call oaxis_check(oa_finebound(h, 0_c_int, level, dtw), "oa_finebound")
      call oaxis_check(oa_auxiliary(h, 0_c_int, level), "oa_auxiliary")
    end do
    call oaxis_check(oa_advance_time(h, 0_c_int, level), "oa_advance_time")       ! :351-357
    if (level < nlev - 1) then                                                     ! :375-378
      call onestep(level + 1_c_int)
      call onestep(level + 1_c_int)
    end if
    call oaxis_check(oa_get_time(h, 0_c_int, level, t, t1, t2, s), "oa_get_time")
    if (level > 0 .and. mod(s, 2) == 0) then                                       ! :421-463
      call oaxis_check(oa_restrict(h, 0_c_int, level), "oa_restrict")
      call oaxis_check(oa_auxiliary(h, 0_c_int, level - 1_c_int), "oa_auxiliary")
    end if
  end subroutine onestep

end module onestep_body

program oaxis_onestep_driver
  use, intrinsic :: iso_c_binding
  use oaxis_mod
  use onestep_body
  implicit none
  type(oaxis_params), target :: par
  character(len=1024) :: parfile, arg, outfile, name
  character(len=32) :: integ
  character(kind=c_char) :: msg(512)
  integer(c_int) :: nsteps, level, Nr, Nz, ghost, s, icn_iter
  real(c_double) :: dr, dz, dt, t, t1, t2
  real(c_double), allocatable, target :: a(:,:)
  integer :: u, step, q, nnames
  character(kind=c_char) :: buf(64)

  call get_command_argument(1, parfile)
  call get_command_argument(2, arg)
  read (arg, *) nsteps
  call get_command_argument(3, outfile)
  nnames = command_argument_count() - 3

  call oaxis_params_init(par)
  call oaxis_check(oa_params_parse_file(c_loc(par%bytes), c_str(parfile)), "oa_params_parse_file")
  call oaxis_check(oa_params_check_supported(c_loc(par%bytes), msg, 512_c_int), "oa_params_check_supported")
  call oaxis_check(oa_params_get(c_loc(par%bytes), c_str("integrator"), buf, 64_c_int), "oa_params_get")
  integ = c_to_text(buf)
  rk4 = (trim(integ) == "rk4")
  call oaxis_check(oa_params_get(c_loc(par%bytes), c_str("icniter"), buf, 64_c_int), "oa_params_get")
  read (c_to_text(buf), *) icn_iter
  niter = merge(4_c_int, icn_iter, rk4)

  h = c_null_ptr
  call oaxis_check(oa_create(c_loc(par%bytes), h), "oa_create")
  call oaxis_check(oa_num_levels(h, nlev), "oa_num_levels")
  call oaxis_check(oa_initial_data(h), "oa_initial_data")
  call oaxis_check(oa_initial_aux(h), "oa_initial_aux")
  do step = 1, nsteps
    call onestep(0_c_int)
  end do

  open(newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
  do level = 0, nlev - 1
    call oaxis_check(oa_level_info(h, 0_c_int, level, Nr, Nz, ghost, dr, dz, dt), "oa_level_info")
    allocate(a(1-ghost:Nr, 1-ghost:Nz))
    do q = 1, nnames
      call get_command_argument(3 + q, name)
      call oaxis_check(oa_download_field(h, 0_c_int, level, c_str(name), a, Nr + ghost), "oa_download_field")
      write (u) a
    end do
    deallocate(a)
    call oaxis_check(oa_get_time(h, 0_c_int, level, t, t1, t2, s), "oa_get_time")
    print '(a,i0,a,i0,a,es24.16,a,es24.16,a,es24.16)', 'level ', level, ' s ', s, ' t ', t, ' t1 ', t1, ' t2 ', t2
  end do
  close(u)
  call oaxis_check(oa_destroy(h), "oa_destroy")

contains

  function c_to_text(b) result(text)
    character(kind=c_char), intent(in) :: b(:)
    character(len=64) :: text
    integer :: i
    text = ' '
    do i = 1, size(b)
      if (b(i) == c_null_char) exit
      text(i:i) = b(i)
    end do
  end function c_to_text

end program oaxis_onestep_driver
