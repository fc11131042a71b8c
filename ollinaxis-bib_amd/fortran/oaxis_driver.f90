! Minimal Fortran main loop over liboaxis.so: what main.f90 + evolve.f90 of
! the reference do for the evolution, with the substep body replaced by the
! C ABI.  Usage:  oaxis_driver <parfile> <coarse steps> <output file> [<checkpoint directory>]
! With a checkpoint directory (and idata = checkpoint in the par file) the
! state comes from the reference's checkpoint files, as initial.f90:157-161
! calls checkpointrestart, instead of the initial data.
! Writes alpha, chi and trK of level 0 as raw doubles in the reference
! layout (1-ghost:Nr, 1-ghost:Nz), column-major, prints a summary line, then
! looks for the apparent horizon at the final time (horizon_finder, as
! evolve.f90:429-432 calls it) and prints what ahshoot reports.
program oaxis_driver
  use, intrinsic :: iso_c_binding
  use oaxis_mod
  implicit none
  type(oaxis_params), target :: par
  type(c_ptr) :: h
  character(len=1024) :: parfile, arg, outfile, ckdir
  integer(c_int) :: nsteps, Nr, Nz, ghost, nlev
  real(c_double) :: dr, dz, dt, t, t1, t2, ms
  integer(c_int) :: s
  real(c_double), allocatable, target :: alpha(:,:), chi(:,:), trK(:,:)
  character(kind=c_char) :: msg(512)
  type(oa_horizon) :: ah
  integer :: u

  call get_command_argument(1, parfile)
  call get_command_argument(2, arg)
  read (arg, *) nsteps
  call get_command_argument(3, outfile)
  ckdir = ''
  if (command_argument_count() >= 4) call get_command_argument(4, ckdir)

  call oaxis_params_init(par)
  call oaxis_check(oa_params_parse_file(c_loc(par%bytes), c_str(parfile)), "oa_params_parse_file")
  call oaxis_check(oa_params_check_supported(c_loc(par%bytes), msg, 512_c_int), "oa_params_check_supported")

  h = c_null_ptr
  call oaxis_check(oa_create(c_loc(par%bytes), h), "oa_create")
  call oaxis_check(oa_num_levels(h, nlev), "oa_num_levels")
  call oaxis_check(oa_level_info(h, 0_c_int, 0_c_int, Nr, Nz, ghost, dr, dz, dt), "oa_level_info")

  if (len_trim(ckdir) > 0) then
    call oaxis_check(oa_checkpoint_restart(h, c_str(ckdir)), "oa_checkpoint_restart")
  else
    call oaxis_check(oa_initial_data(h), "oa_initial_data")
  end if
  call oaxis_check(oa_initial_aux(h), "oa_initial_aux")
  call oaxis_check(oa_evolve(h, nsteps), "oa_evolve")
  call oaxis_check(oa_last_elapsed_ms(h, ms), "oa_last_elapsed_ms")
  call oaxis_check(oa_get_time(h, 0_c_int, 0_c_int, t, t1, t2, s), "oa_get_time")

  allocate(alpha(1-ghost:Nr, 1-ghost:Nz), chi(1-ghost:Nr, 1-ghost:Nz), trK(1-ghost:Nr, 1-ghost:Nz))
  call oaxis_check(oa_download_field(h, 0_c_int, 0_c_int, c_str("alpha"), alpha, Nr + ghost), "download alpha")
  call oaxis_check(oa_download_field(h, 0_c_int, 0_c_int, c_str("chi"), chi, Nr + ghost), "download chi")
  call oaxis_check(oa_download_field(h, 0_c_int, 0_c_int, c_str("trK"), trK, Nr + ghost), "download trK")

  open(newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
  write (u) alpha, chi, trK
  close(u)
  print '(a,i0,a,i0,a,i0,a,i0,a,es24.16,a,es24.16,a,f10.3)', 'levels ', nlev, ' Nr ', Nr, ' Nz ', Nz, &
      ' steps ', s, ' t ', t, ' min(alpha) ', minval(alpha), ' ms ', ms
  call oaxis_check(oa_ahfind(h, 0_c_int, ah), "oa_ahfind")
  print '(a,i0,a,i0,a,es24.16,a,es24.16,a,es24.16)', 'horizon found ', ah%found, ' shots ', ah%shots, &
      ' zinit ', ah%zinit, ' area ', ah%area, ' mass ', ah%mass
  call oaxis_check(oa_destroy(h), "oa_destroy")
end program oaxis_driver
