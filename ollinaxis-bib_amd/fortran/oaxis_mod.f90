! ISO_C_BINDING interface to liboaxis.so, the MI355X evolution path of
! OllinAxis-BiB.  A Fortran main loop (main.f90 / evolve.f90 / onestep.f90 of
! the reference) calls these in place of the routines each one names.
!
! oa_params is handled as an opaque buffer of oa_params_size() bytes: it is
! filled from the reference's own par file (oa_params_parse_file) or key by
! key (oa_params_set), so the Fortran side never mirrors its layout.
! Every function returns 0 (OA_OK) on success; oaxis_check turns a nonzero
! code into the reference's fatal-error behaviour (print + stop, die.f90).
module oaxis_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: OA_OK = 0

  public :: oaxis_params, oaxis_params_init, oaxis_check, c_str
  public :: oa_params_default, oa_params_set, oa_params_parse_file, oa_params_get, &
            oa_params_check_supported, oa_params_size, oa_last_error, &
            oa_create, oa_destroy, oa_num_levels, oa_level_info, &
            oa_upload_field, oa_download_field, oa_get_time, oa_set_time, &
            oa_saveold, oa_substep, oa_finebound, oa_auxiliary, oa_restrict, oa_advance_time, &
            oa_initial_aux, oa_coarse_step, oa_evolve, oa_sync, oa_initial_data, &
            oa_last_elapsed_ms, oa_point_updates, oa_params_parse_text, oa_set_kernel_timing, &
            oa_rhs_kernel_stats, oa_norms0d, oa_norms0d_noalloc, oa_params_footprint, &
            oa_set_elliptic_exit, oa_elliptic_stats, &
            oa_batch_create, oa_batch_destroy, oa_batch_add, oa_batch_remove, oa_batch_size, &
            oa_batch_coarse_step, oa_batch_evolve, oa_batch_sync, oa_batch_last_elapsed_ms, &
            oa_batch_point_updates, oa_batch_launch_stats, oa_batch_set_kernel_timing, &
            oa_batch_rhs_kernel_stats, oa_set_graphs, oa_batch_set_graphs, oa_schedule, oa_batch_schedule, &
            oa_constraints, oa_ahfind, oa_horizon_shape, oa_batch_ahfind, &
            oa_ahfind_begin, oa_batch_ahfind_begin, oa_ahfind_end, &
            oa_checkpoint_size, oa_checkpoint_save, oa_checkpoint_load, oa_checkpoint_restart, &
            oa_checkpoint_write

  ! result of one apparent-horizon search (oa_horizon in include/oaxis.h)
  type, bind(C), public :: oa_horizon
    integer(c_int) :: found, ntheta, shots, errorflag
    real(c_double) :: zcenter, zinit, dradlast, area, radius, mass
    integer(c_int) :: method, iterations
    real(c_double) :: residual
  end type oa_horizon

  ! storage for one oa_params block
  type :: oaxis_params
    integer(c_int8_t), allocatable :: bytes(:)
  end type oaxis_params

  interface
    ! include/oaxis.h: parameter block (param.f90, parse.f90)
    integer(c_int) function oa_params_default(p) bind(C, name="oa_params_default")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    integer(c_int) function oa_params_set(p, key, val) bind(C, name="oa_params_set")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: key, val
    end function
    integer(c_int) function oa_params_parse_file(p, path) bind(C, name="oa_params_parse_file")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: path
    end function
    integer(c_int) function oa_params_get(p, key, out, outlen) bind(C, name="oa_params_get")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: key
      character(kind=c_char), dimension(*) :: out
      integer(c_int), value :: outlen
    end function
    integer(c_int) function oa_params_check_supported(p, msg, msglen) bind(C, name="oa_params_check_supported")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: msg
      integer(c_int), value :: msglen
    end function
    integer(c_size_t) function oa_params_size() bind(C, name="oa_params_size")
      import :: c_size_t
    end function
    type(c_ptr) function oa_last_error() bind(C, name="oa_last_error")
      import :: c_ptr
    end function

    ! handles (main.f90 allocatearrays / setupboxes.f90)
    integer(c_int) function oa_create(p, h) bind(C, name="oa_create")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      type(c_ptr) :: h
    end function
    integer(c_int) function oa_destroy(h) bind(C, name="oa_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function oa_num_levels(h, n) bind(C, name="oa_num_levels")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: n
    end function
    integer(c_int) function oa_level_info(h, box, level, Nr, Nz, ghost, dr, dz, dt) &
        bind(C, name="oa_level_info")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      integer(c_int) :: Nr, Nz, ghost
      real(c_double) :: dr, dz, dt
    end function

    ! grid(box,level)%var <-> device, reference layout (1-ghost:Nr, 1-ghost:Nz)
    integer(c_int) function oa_upload_field(h, box, level, name, src, ld) bind(C, name="oa_upload_field")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      character(kind=c_char), dimension(*) :: name
      real(c_double), dimension(*) :: src
      integer(c_int), value :: ld
    end function
    integer(c_int) function oa_download_field(h, box, level, name, dst, ld) bind(C, name="oa_download_field")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      character(kind=c_char), dimension(*) :: name
      real(c_double), dimension(*) :: dst
      integer(c_int), value :: ld
    end function
    integer(c_int) function oa_get_time(h, box, level, t, t1, t2, s) bind(C, name="oa_get_time")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      real(c_double) :: t, t1, t2
      integer(c_int) :: s
    end function
    integer(c_int) function oa_set_time(h, box, level, t, t1, t2, s) bind(C, name="oa_set_time")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      real(c_double), value :: t, t1, t2
      integer(c_int), value :: s
    end function

    ! the substep body of onestep.f90, callable piecewise
    integer(c_int) function oa_saveold(h, box, level) bind(C, name="oa_saveold")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
    end function
    integer(c_int) function oa_substep(h, box, level, iter, niter, dtw, w) bind(C, name="oa_substep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level, iter, niter
      real(c_double), value :: dtw, w
    end function
    integer(c_int) function oa_finebound(h, box, level, dtw) bind(C, name="oa_finebound")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      real(c_double), value :: dtw
    end function
    integer(c_int) function oa_auxiliary(h, box, level) bind(C, name="oa_auxiliary")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
    end function
    integer(c_int) function oa_restrict(h, box, level) bind(C, name="oa_restrict")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
    end function
    integer(c_int) function oa_advance_time(h, box, level) bind(C, name="oa_advance_time")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
    end function

    ! whole steps (evolve.f90 / onestep(0))
    integer(c_int) function oa_initial_data(h) bind(C, name="oa_initial_data")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function oa_initial_aux(h) bind(C, name="oa_initial_aux")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function oa_coarse_step(h) bind(C, name="oa_coarse_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function oa_evolve(h, n) bind(C, name="oa_evolve")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n
    end function
    integer(c_int) function oa_sync(h) bind(C, name="oa_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function oa_last_elapsed_ms(h, ms) bind(C, name="oa_last_elapsed_ms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: ms
    end function
    integer(c_int) function oa_point_updates(h, u) bind(C, name="oa_point_updates")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: u
    end function
    integer(c_int) function oa_params_parse_text(p, text) bind(C, name="oa_params_parse_text")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: p
      character(kind=c_char), dimension(*) :: text
    end function
    integer(c_int) function oa_set_graphs(h, on) bind(C, name="oa_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function oa_batch_set_graphs(b, on) bind(C, name="oa_batch_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: on
    end function
    integer(c_int) function oa_schedule(h, graphs, xcd_remap, tuned, step_ms) bind(C, name="oa_schedule")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int) :: graphs, xcd_remap, tuned
      real(c_double), dimension(4) :: step_ms
    end function
    integer(c_int) function oa_batch_schedule(b, graphs, xcd_remap, tuned, step_ms) &
        bind(C, name="oa_batch_schedule")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int) :: graphs, xcd_remap, tuned
      real(c_double), dimension(4) :: step_ms
    end function
    integer(c_int) function oa_set_kernel_timing(h, on) bind(C, name="oa_set_kernel_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function oa_rhs_kernel_stats(h, avg_ms, launches) bind(C, name="oa_rhs_kernel_stats")
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: h
      real(c_double) :: avg_ms
      integer(c_long) :: launches
    end function
    ! constraints.f90 on a level's current state (ham, mom_r, mom_z, CDelta_r, CDelta_z)
    integer(c_int) function oa_constraints(h, box, level) bind(C, name="oa_constraints")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
    end function
    ! checkpointsave / checkpointload: the whole state as one byte block
    integer(c_int) function oa_checkpoint_size(h, bytes) bind(C, name="oa_checkpoint_size")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h
      integer(c_size_t) :: bytes
    end function
    integer(c_int) function oa_checkpoint_save(h, buf, bytes) bind(C, name="oa_checkpoint_save")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h, buf
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function oa_checkpoint_load(h, buf, bytes) bind(C, name="oa_checkpoint_load")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h, buf
      integer(c_size_t), value :: bytes
    end function
    ! checkpointsave: the state as a reference checkpoint directory
    integer(c_int) function oa_checkpoint_write(h, dir, parfile) bind(C, name="oa_checkpoint_write")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*) :: dir, parfile
    end function
    ! checkpointrestart: the state from a reference checkpoint directory
    integer(c_int) function oa_checkpoint_restart(h, dir) bind(C, name="oa_checkpoint_restart")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*) :: dir
    end function
    ! horizon_finder for horizon k on the current state (par file's ahmethod)
    integer(c_int) function oa_ahfind(h, k, res) bind(C, name="oa_ahfind")
      import :: c_int, c_ptr, oa_horizon
      type(c_ptr), value :: h
      integer(c_int), value :: k
      type(oa_horizon) :: res
    end function
    ! (theta, h(theta)) of the last integration of horizon k (ah_radius_SH)
    integer(c_int) function oa_horizon_shape(h, k, theta, rad, cap, n) bind(C, name="oa_horizon_shape")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: k, cap
      real(c_double), dimension(*) :: theta, rad
      integer(c_int) :: n
    end function
    ! horizon_finder for horizon k(i) of members(i), i <= n, in lockstep (no reference counterpart)
    integer(c_int) function oa_batch_ahfind(b, n, members, k, res) bind(C, name="oa_batch_ahfind")
      import :: c_int, c_ptr, oa_horizon
      type(c_ptr), value :: b
      integer(c_int), value :: n
      type(c_ptr), dimension(*) :: members
      integer(c_int), dimension(*) :: k
      type(oa_horizon), dimension(*) :: res
    end function
    ! a horizon search in two calls: the begin reads the current state, the
    ! end runs the shots beside whatever was enqueued in between
    integer(c_int) function oa_ahfind_begin(h, k, search) bind(C, name="oa_ahfind_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: k
      type(c_ptr) :: search
    end function
    integer(c_int) function oa_batch_ahfind_begin(b, n, members, k, search) bind(C, name="oa_batch_ahfind_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: n
      type(c_ptr), dimension(*) :: members
      integer(c_int), dimension(*) :: k
      type(c_ptr) :: search
    end function
    integer(c_int) function oa_ahfind_end(search, res) bind(C, name="oa_ahfind_end")
      import :: c_int, c_ptr, oa_horizon
      type(c_ptr), value :: search
      type(oa_horizon), dimension(*) :: res
    end function
    ! 0D norms (max, min, nm1, nm2) of an evolved array, as save0Ddata forms them
    integer(c_int) function oa_norms0d(h, box, level, name, norms) bind(C, name="oa_norms0d")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: h
      integer(c_int), value :: box, level
      character(kind=c_char), dimension(*) :: name
      real(c_double), dimension(4) :: norms
    end function
    ! the same norms of n arrays into device memory (5 doubles each), with no
    ! allocation: names is an array of n pointers to C strings (c_loc of each)
    integer(c_int) function oa_norms0d_noalloc(h, box, level, n, names, out_dev) bind(C, name="oa_norms0d_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: box, level, n
      type(c_ptr), dimension(*) :: names
      type(c_ptr), value :: out_dev
    end function
    ! stopping rule of the initial data's elliptic solve (1: stationary exit too)
    integer(c_int) function oa_set_elliptic_exit(h, stationary) bind(C, name="oa_set_elliptic_exit")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: stationary
    end function
    integer(c_int) function oa_elliptic_stats(h, steps0, steps1, residual) bind(C, name="oa_elliptic_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int) :: steps0, steps1
      real(c_double) :: residual
    end function
    integer(c_int) function oa_params_footprint(p, bytes) bind(C, name="oa_params_footprint")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: p
      real(c_double) :: bytes
    end function
    ! batches of spacetimes (no counterpart in the reference)
    integer(c_int) function oa_batch_create(b) bind(C, name="oa_batch_create")
      import :: c_int, c_ptr
      type(c_ptr) :: b
    end function
    integer(c_int) function oa_batch_destroy(b) bind(C, name="oa_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function oa_batch_add(b, h) bind(C, name="oa_batch_add")
      import :: c_int, c_ptr
      type(c_ptr), value :: b, h
    end function
    integer(c_int) function oa_batch_remove(b, h) bind(C, name="oa_batch_remove")
      import :: c_int, c_ptr
      type(c_ptr), value :: b, h
    end function
    integer(c_int) function oa_batch_size(b, n) bind(C, name="oa_batch_size")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: n
    end function
    integer(c_int) function oa_batch_coarse_step(b) bind(C, name="oa_batch_coarse_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function oa_batch_evolve(b, nsteps) bind(C, name="oa_batch_evolve")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function oa_batch_sync(b) bind(C, name="oa_batch_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function oa_batch_last_elapsed_ms(b, ms) bind(C, name="oa_batch_last_elapsed_ms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      real(c_double) :: ms
    end function
    integer(c_int) function oa_batch_point_updates(b, u) bind(C, name="oa_batch_point_updates")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      real(c_double) :: u
    end function
    integer(c_int) function oa_batch_launch_stats(b, launches, items) bind(C, name="oa_batch_launch_stats")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: b
      integer(c_long) :: launches, items
    end function
    integer(c_int) function oa_batch_set_kernel_timing(b, on) bind(C, name="oa_batch_set_kernel_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: on
    end function
    integer(c_int) function oa_batch_rhs_kernel_stats(b, total_ms, launches, points) &
        bind(C, name="oa_batch_rhs_kernel_stats")
      import :: c_int, c_ptr, c_double, c_long
      type(c_ptr), value :: b
      real(c_double) :: total_ms, points
      integer(c_long) :: launches
    end function
  end interface

contains

  ! Fortran string -> NUL-terminated C string
  function c_str(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function

  ! allocate a parameter block and fill it with the reference defaults
  subroutine oaxis_params_init(p)
    type(oaxis_params), target, intent(inout) :: p
    if (allocated(p%bytes)) deallocate(p%bytes)
    allocate(p%bytes(oa_params_size()))
    call oaxis_check(oa_params_default(c_loc(p%bytes)), "oa_params_default")
  end subroutine

  ! nonzero return code -> message and stop, as die.f90 does
  subroutine oaxis_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == OA_OK) return
    call c_f_pointer(oa_last_error(), msg, [4096])
    n = 0
    do while (n < 4096)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    print *, trim(what), ' failed (code ', rc, '): ', transfer(msg(1:n), repeat(' ', n))
    stop 1
  end subroutine

end module oaxis_mod
