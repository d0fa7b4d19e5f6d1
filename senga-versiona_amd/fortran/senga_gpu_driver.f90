! senga_gpu_driver.f90 -- SENGA+ time loop on one MI355X.
!
! The loop of src/senga2.f:96-194 with the substage body on the GPU:
!   read input/cont.dat, input/chem.dat, input/diff.dat (src/indata.f, src/contin.f,
!   src/chemin.f, src/diffin.f); restart from output/dmpi0000000.dat
!   (src/indata.f:1419-1475); for each step five sg_substage calls in place of
!   BOUNDT..LINCOM/FINCOM and sg_end_step in place of BOUNTT; dumps every NTDUMP
!   steps, alternating output/dmpi0000001.dat and output/dmpi0000000.dat (the
!   first dump does not overwrite the restart file) as src/output.f:273-335 does.
!
! The run writes output/rept.res, output/stat.res and, on every dump step,
! output/outIIIIIPPPP.res as src/output.f:81-99,241-265,343-352,397-425 do.
!
! Usage:  senga_gpu [rundir]          run the case in rundir (default '.')
!         senga_gpu --config rundir out   write the sg_config built from rundir's
!                                         inputs to file out and stop (no GPU use)
!         senga_gpu --convert-dump in out  read a formatted dump and write it
!                                         unformatted (no GPU use)
program senga_gpu_driver
  use, intrinsic :: iso_c_binding
  use senga_gpu
  use senga_io
  implicit none
  type(control_t) :: ctl
  type(sg_config_t), target :: cfg
  type(c_ptr) :: h
  character(len=512) :: rundir, arg, outcfg
  real(c_double), allocatable, target :: run(:, :), tfld(:), pfld(:), rate(:, :)
  real(c_double) :: etime, tstep, errold, errldr, t0, t1
  integer :: ierr, nargs, itime, irk, ns, npts, idflag, u
  integer(c_int) :: rc, compiled
  integer(c_int8_t), pointer :: bytes(:)
  logical :: cfgonly
  integer(8) :: c0, c1, crate
  integer :: dnx, dny, dnz, dns
  character(len=32) :: fname

  nargs = command_argument_count()
  cfgonly = .false.
  rundir = '.'
  if (nargs >= 1) then
    call get_command_argument(1, arg)
    if (trim(arg) == '--config') then
      cfgonly = .true.
      call get_command_argument(2, rundir)
      call get_command_argument(3, outcfg)
    else if (trim(arg) == '--convert-dump') then
      call get_command_argument(2, rundir)
      call get_command_argument(3, outcfg)
      call dump_header(trim(rundir), .true., dnx, dny, dnz, dns, ierr)
      if (ierr /= 0) call fail('cannot read the formatted dump header')
      allocate(run(dnx * dny * dnz, 5 + dns))
      call read_dump(trim(rundir), .true., dnx, dny, dnz, dns, run, etime, tstep, errold, errldr, ierr)
      if (ierr /= 0) call fail('cannot read the formatted dump')
      call write_dump(trim(outcfg), .false., dnx, dny, dnz, dns, run, etime, tstep, errold, errldr, ierr)
      if (ierr /= 0) call fail('cannot write the unformatted dump')
      stop
    else
      rundir = arg
    end if
  end if

  call zero_config(cfg)
  call read_cont(trim(rundir) // '/input/cont.dat', ctl, ierr)
  if (ierr /= 0) call fail('cannot read input/cont.dat')
  call read_chem(trim(rundir) // '/input/chem.dat', cfg, ierr)
  if (ierr /= 0) call fail('cannot read input/chem.dat')
  if (ctl%nfmavt == 1) then
    call read_diff(trim(rundir) // '/input/diff.dat', cfg, ierr)
    if (ierr /= 0) call fail('cannot read input/diff.dat')
  end if
  if (ctl%nfradn == 1) then
    call read_radn(trim(rundir) // '/input/radn.dat', cfg, ierr)
    if (ierr /= 0) call fail('cannot read input/radn.dat')
  end if
  call fill_config(ctl, cfg)

  if (cfgonly) then
    call c_f_pointer(c_loc(cfg), bytes, [c_sizeof(cfg)])
    open(newunit=u, file=trim(outcfg), status='replace', access='stream', form='unformatted')
    write(u) bytes
    close(u)
  else
    call run_case()
  end if

contains

  subroutine run_case()

  if (sg_abi_version() /= 4 .or. sg_config_size() /= c_sizeof(cfg)) call fail('library ABI mismatch')
  ns = cfg%nspec
  npts = ctl%nx * ctl%ny * ctl%nz
  allocate(run(npts, 5 + ns))
  if (ctl%ncdmpi /= 1) call fail('cold start is not supported: provide output/dmpi0000000.dat')
  call read_dump(trim(rundir) // '/output/dmpi0000000.dat', ctl%ndifmt == 1, ctl%nx, ctl%ny, ctl%nz, ns, &
                 run, etime, tstep, errold, errldr, ierr)
  if (ierr /= 0) call fail('cannot read output/dmpi0000000.dat')
  ! On restart with a fixed step the reference keeps cont.dat's TSTEP through
  ! the REAL variable TSOLD, i.e. rounded to single precision; with the
  ! adaptive step it keeps the dump's TSTEP (src/indata.f:1423,1481).
  if (ctl%nstpsw == 0) tstep = real(real(ctl%tstep, c_float), c_double)

  rc = sg_create(cfg, h)
  if (rc /= SG_OK) call fail('sg_create failed (unsupported configuration or no device)')
  rc = sg_set_timestep(h, tstep)
  if (ctl%nstpsw == 1) then
    rc = sg_set_adaptive(h, 1_c_int, errold, errldr)   ! ADAPTT after each step
    if (rc /= SG_OK) call sgfail(h, rc)
  end if
  rc = sg_upload_state(h, run, etime)
  if (rc /= SG_OK) call sgfail(h, rc)

  allocate(tfld(npts), pfld(npts), rate(npts, ns))
  call report_open(trim(rundir) // '/output/rept.res', ctl, ns, tstep)
  call stat_open(trim(rundir) // '/output/stat.res')
  write(*, '(a, 3i6, a, i3, a, es12.5)') 'senga_gpu: grid', ctl%nx, ctl%ny, ctl%nz, ' species', ns, ' tstep', tstep
  rc = sg_mechanism_compiled(h, compiled)
  if (compiled /= 0) then
    write(*, '(a)') 'senga_gpu: chemistry runs the compiled mechanism (unrolled step loop)'
  else
    write(*, '(a)') 'senga_gpu: chemistry runs the table-driven step loop'
  end if
  idflag = 1   ! the first dump goes to dmpi0000001.dat (src/indata.f:127-134 leaves IDFLAG = 1)
  call system_clock(c0, crate)
  do itime = ctl%ntime1, ctl%ntime1 + ctl%ntime - 1
    do irk = 1, 5
      rc = sg_substage(h, int(irk, c_int))         ! BOUNDT .. LINCOM / FINCOM
      if (rc /= SG_OK) call sgfail(h, rc)
    end do
    rc = sg_end_step(h)                             ! ETIME += TSTEP; BOUNTT; ADAPTT
    if (rc /= SG_OK) call sgfail(h, rc)
    etime = etime + tstep
    rc = sg_get_timestep(h, tstep, errold, errldr)
    ! OUTPUT (src/output.f): report, post-processing file and dump, statistics
    if (mod(itime, ctl%ntrept) == 0) call report_step(trim(rundir) // '/output/rept.res', itime, etime, tstep)
    if (mod(itime, ctl%ntdump) == 0) then
      rc = sg_download_state(h, c_loc(run), c_loc(tfld), c_loc(pfld), c_loc(rate))
      if (rc /= SG_OK) call sgfail(h, rc)
      write(fname, '(a, i5.5, i4.4, a)') 'output/out', itime / ctl%ntdump, 0, '.res'
      call write_out_res(trim(rundir) // '/' // trim(fname), ctl%nx, ctl%ny, ctl%nz, ns, run, tfld, pfld, &
                         rate, etime, ierr)
      if (ierr /= 0) call fail('cannot write ' // trim(fname))
      write(fname, '(a, i1, a)') 'output/dmpi000000', idflag, '.dat'
      call write_dump(trim(rundir) // '/' // trim(fname), ctl%ndofmt == 1, ctl%nx, ctl%ny, ctl%nz, ns, run, &
                      etime, tstep, errold, errldr, ierr)
      if (ierr /= 0) call fail('cannot write dump')
      call report_dump(trim(rundir) // '/output/rept.res', trim(fname))
      idflag = 1 - idflag
    end if
    if (ctl%ntstat > 0) then
      if (mod(itime, ctl%ntstat) == 0) call stat_step(trim(rundir) // '/output/stat.res', itime)
    end if
  end do
  rc = sg_sync(h)
  if (rc /= SG_OK) call sgfail(h, rc)
  call system_clock(c1)
  t0 = 0.0d0
  t1 = real(c1 - c0, c_double) / real(crate, c_double)
  if (t1 > 0.0d0) write(*, '(a, i6, a, f10.3, a, f12.3, a)') 'senga_gpu: ', ctl%ntime, ' steps in ', t1, &
      ' s, ', real(npts, c_double) * 5.0d0 * ctl%ntime / t1 / 1.0d6, ' Mpoint-RK-substages/s (with dumps)'
  rc = sg_destroy(h)
  end subroutine run_case

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a)') 'Fatal: senga_gpu: ' // msg
    error stop 1
  end subroutine fail

  ! The reference's fatal-error behaviour (src/temper.f:156-165): message,
  ! failing point, stop.
  subroutine sgfail(h, rc)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: rc
    integer(c_int) :: code, ijk(3), r
    character(kind=c_char) :: msg(256)
    character(len=256) :: text
    integer :: k
    r = sg_last_error(h, code, ijk, msg, 256_c_size_t)
    text = ''
    do k = 1, 256
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write(*, '(a, i0, a, a)') 'Fatal: error ', rc, ': ', trim(text)
    write(*, '(a, 3i6)') 'at point:', ijk
    error stop 2
  end subroutine sgfail

end program senga_gpu_driver
