! senga_io.f90 -- the SENGA+ input and dump formats for the GPU driver.
!
! Readers for input/cont.dat (layout of src/contin.f:83-214), input/chem.dat
! (src/chemin.f:155-344) and input/diff.dat (src/diffin.f:95-212), the derived
! thermochemical tables of src/chemin.f:822-983 and the transport tables of
! src/diffin.f:280-405, written into the C-ABI configuration sg_config_t; and
! the restart dump dmpi in its unformatted and formatted forms
! (src/output.f:306-333, src/indata.f:1430-1475).  The edit descriptors are
! those of the file formats, so files written for the reference read here
! unchanged.  The run's outputs follow src/output.f: the report file
! output/rept.res (time step, elapsed time and next step every NTREPT steps,
! one line per dump), the statistics file output/stat.res (the step number
! every NTSTAT steps) and the post-processing file output/outIIIIIPPPP.res of
! every dump step.
module senga_io
  use, intrinsic :: iso_c_binding
  use senga_gpu
  implicit none
  private

  ! physical constants of src/com_senga2.h:112-119
  real(c_double), parameter :: RGUNIV = 8.3142d3
  real(c_double), parameter :: ALAMDC = 2.58d-5, RLAMDA = 0.7d0, TLAMDA = 298.0d0
  real(c_double), parameter :: PRANTL = 0.7d0

  ! run control (cont.dat)
  type, public :: control_t
    real(c_double) :: xgdlen, ygdlen, zgdlen
    integer :: nx, ny, nz
    real(c_double) :: tstep
    integer :: ntime1, ntime, nstpsw
    integer :: ntdump, ntrept, ntstat, ndofmt
    integer :: ncdmpi, ndifmt
    real(c_double) :: prin, trin, urin, vrin, wrin
    integer :: nyrin
    real(c_double) :: yrin(SG_MAX_SPEC)
    integer :: bctype(6), bcipar(4, 6)
    real(c_double) :: bcrpar(4, 6)
    ! molecular transport and radiation switches (src/contin.f:194-209)
    integer :: nfmavt, nfmixw, nfmixp, nfmsor, nfmduf, nfradn
    real(c_double) :: wmltdr, trefrn
  end type control_t

  public :: read_cont, read_chem, read_diff, read_radn, fill_config, zero_config
  public :: read_dump, write_dump, dump_header
  public :: report_open, report_step, report_dump, stat_open, stat_step, write_out_res

contains

  subroutine zero_config(cfg)
    type(sg_config_t), target, intent(inout) :: cfg
    integer(c_int8_t), pointer :: b(:)
    call c_f_pointer(c_loc(cfg), b, [c_sizeof(cfg)])
    b = 0_c_int8_t
  end subroutine zero_config

  ! ---------------------------------------------------------------- cont.dat
  subroutine read_cont(path, ctl, ierr)
    character(len=*), intent(in) :: path
    type(control_t), intent(out) :: ctl
    integer, intent(out) :: ierr
    integer :: u, i, j, nprx, npry, nprz, inturb, inseed, inflam
    real(c_double) :: sparam(4)
    character(len=64) :: line
    ctl%yrin = 0.0d0
    ! defaults of src/contin.f:194-201 when the block is absent
    ctl%nfmavt = 0
    ctl%nfmixw = 0
    ctl%nfmixp = 0
    ctl%nfmsor = 0
    ctl%nfmduf = 0
    ctl%wmltdr = 0.0d0
    ctl%nfradn = 0
    ctl%trefrn = 3.0d0
    open(newunit=u, file=path, status='old', form='formatted', iostat=ierr)
    if (ierr /= 0) return
    do i = 1, 7
      read(u, *)
    end do
    read(u, *) ctl%xgdlen, ctl%ygdlen, ctl%zgdlen
    read(u, *)
    read(u, *)
    read(u, *) ctl%nx, ctl%ny, ctl%nz
    read(u, *)
    read(u, *)
    read(u, *) nprx, npry, nprz
    read(u, *)
    read(u, *)
    read(u, *) ctl%tstep, ctl%ntime1, ctl%ntime, ctl%nstpsw
    read(u, *)
    read(u, *)
    read(u, *) ctl%ntdump, ctl%ntrept, ctl%ntstat, ctl%ndofmt
    read(u, *)
    read(u, *)
    read(u, *) ctl%ncdmpi, ctl%ndifmt
    read(u, *)
    read(u, *)
    read(u, *)
    read(u, *) inturb, inseed, sparam
    read(u, *)
    read(u, *)
    read(u, *) inflam
    read(u, *)
    read(u, *)
    read(u, *)
    read(u, *) ctl%prin, ctl%trin, ctl%urin, ctl%vrin, ctl%wrin
    read(u, *)
    read(u, *) ctl%nyrin
    if (ctl%nyrin > SG_MAX_SPEC) then
      ierr = 1
      close(u)
      return
    end if
    do i = 1, ctl%nyrin
      read(u, *) j, ctl%yrin(i)
    end do
    do i = 1, 5
      read(u, *)
    end do
    do i = 1, 6
      read(u, *) ctl%bctype(i), ctl%bcipar(:, i), ctl%bcrpar(:, i)
    end do
    read(u, *)
    read(u, '(a)', iostat=ierr) line
    if (ierr == 0 .and. line(1:32) == 'Molecular transport control data') then
      ! src/contin.f:203-210: switches, three label lines, radiation switch
      read(u, *)
      read(u, *) ctl%nfmavt, ctl%nfmixw, ctl%nfmixp, ctl%nfmsor, ctl%nfmduf, ctl%wmltdr
      read(u, *)
      read(u, *)
      read(u, *)
      read(u, *) ctl%nfradn, ctl%trefrn
    end if
    ierr = 0
    close(u)
  end subroutine read_cont

  ! ---------------------------------------------------------------- chem.dat
  ! Reads the mechanism and fills the thermochemical members of cfg with the
  ! derived tables (src/chemin.f:822-983), species lists 0-based.
  subroutine read_chem(path, cfg, ierr)
    character(len=*), intent(in) :: path
    type(sg_config_t), intent(inout) :: cfg
    integer, intent(out) :: ierr
    integer :: u, ns, nst, i, j, k, isp, ist, n, it, nc, icp, nbody, ngibb, nlind, ntroe, nsrif
    character(len=32) :: sym
    real(c_double) :: prefgb, giblet, clewis(SG_MAX_SPEC), tlo, thi
    real(c_double) :: amolcp(SG_NCOF, SG_NTINT, SG_MAX_SPEC), diffmu_r, rdat(12)
    integer :: ncofcp(SG_NTINT, SG_MAX_SPEC)
    open(newunit=u, file=path, status='old', form='formatted', iostat=ierr)
    if (ierr /= 0) return
    do i = 1, 6
      read(u, *)
    end do
    read(u, '(I5)') ns
    if (ns < 1 .or. ns > SG_MAX_SPEC) then
      ierr = 1
      close(u)
      return
    end if
    cfg%nspec = ns
    do i = 1, ns
      read(u, '(I5,3X,A)') isp, sym
    end do
    read(u, *)
    read(u, *) prefgb
    amolcp = 0.0d0
    ncofcp = 0
    do i = 1, ns
      read(u, '(I5,1PE12.4)') isp, tlo
      if (isp < 1 .or. isp > ns) goto 900
      cfg%wmolar(isp) = tlo
      read(u, '(1PE12.4)') clewis(isp)
      read(u, '(I5)') n
      if (n < 1 .or. n > SG_NTINT) goto 900
      cfg%ntint(isp) = n
      do it = 1, cfg%ntint(isp)
        read(u, '(2(1PE12.4),I5)') tlo, thi, n
        if (n < 1 .or. n > SG_NCOF) goto 900
        ncofcp(it, isp) = n
        cfg%tinthi(it, isp) = thi
        do icp = 1, ncofcp(it, isp)
          read(u, '(1PE15.7)') amolcp(icp, it, isp)
        end do
      end do
    end do
    read(u, *)
    read(u, '(I5)') nst
    if (nst < 0 .or. nst > SG_MAX_STEP) then
      ierr = 1
      close(u)
      return
    end if
    cfg%nstep = nst
    do j = 1, nst
      read(u, '(I5,3(1PE12.4))') ist, tlo, thi, diffmu_r
      if (ist < 1 .or. ist > nst) goto 900
      cfg%rparam(:, ist) = [tlo, thi, diffmu_r]
    end do
    ! species, reactant and product lists
    read(u, *)
    do j = 1, nst
      read(u, '(2I5)') ist, n
      if (ist < 1 .or. ist > nst .or. n < 0 .or. n > SG_MAX_SSL) goto 900
      cfg%nsslen(ist) = n
      do k = 1, cfg%nsslen(ist)
        read(u, '(2I5,I8)') ist, i, n
        if (ist < 1 .or. ist > nst .or. i < 1 .or. i > SG_MAX_SSL .or. n < 1 .or. n > ns) goto 900
        cfg%nsspec(i, ist) = n - 1
      end do
    end do
    read(u, *)
    do j = 1, nst
      read(u, '(2I5)') ist, n
      if (ist < 1 .or. ist > nst .or. n < 0 .or. n > SG_MAX_SSL) goto 900
      cfg%nrslen(ist) = n
      do k = 1, cfg%nrslen(ist)
        read(u, '(2I5,I8)') ist, i, n
        if (ist < 1 .or. ist > nst .or. i < 1 .or. i > SG_MAX_SSL .or. n < 1 .or. n > ns) goto 900
        cfg%nrspec(i, ist) = n - 1
      end do
    end do
    read(u, *)
    do j = 1, nst
      read(u, '(2I5)') ist, n
      if (ist < 1 .or. ist > nst .or. n < 0 .or. n > SG_MAX_SSL) goto 900
      cfg%npslen(ist) = n
      do k = 1, cfg%npslen(ist)
        read(u, '(2I5,I8)') ist, i, n
        if (ist < 1 .or. ist > nst .or. i < 1 .or. i > SG_MAX_SSL .or. n < 1 .or. n > ns) goto 900
        cfg%npspec(i, ist) = n - 1
      end do
    end do
    read(u, *)
    do j = 1, nst
      read(u, '(2I5)') ist, n
      if (ist < 1 .or. ist > nst .or. n < 0 .or. n > SG_MAX_SSL) goto 900
      cfg%nrclen(ist) = n
      do k = 1, cfg%nrclen(ist)
        read(u, '(2I5,I8,1PE12.4)') ist, i, n, diffmu_r
        if (ist < 1 .or. ist > nst .or. i < 1 .or. i > SG_MAX_SSL .or. n < 1 .or. n > ns) goto 900
        cfg%crspec(i, ist) = diffmu_r
        cfg%nrcpec(i, ist) = n - 1
      end do
    end do
    read(u, *)
    do j = 1, nst
      read(u, '(2I5)') ist, n
      if (ist < 1 .or. ist > nst .or. n < 0 .or. n > SG_MAX_SSL) goto 900
      cfg%npclen(ist) = n
      do k = 1, cfg%npclen(ist)
        read(u, '(2I5,I8,1PE12.4)') ist, i, n, diffmu_r
        if (ist < 1 .or. ist > nst .or. i < 1 .or. i > SG_MAX_SSL .or. n < 1 .or. n > ns) goto 900
        cfg%cpspec(i, ist) = diffmu_r
        cfg%npcpec(i, ist) = n - 1
      end do
    end do
    read(u, *)
    do j = 1, nst
      do k = 1, cfg%nsslen(j)
        read(u, '(2I5,F5.1)') ist, i, diffmu_r
        if (ist < 1 .or. ist > nst .or. i < 1 .or. i > SG_MAX_SSL) goto 900
        cfg%diffmu(i, ist) = diffmu_r
      end do
    end do
    ! third bodies
    read(u, *)
    read(u, '(I5)') nbody
    if (nbody < 0 .or. nbody > SG_MAX_BODY) goto 900
    do i = 1, nbody
      read(u, '(I5,3X,A)') n, sym
    end do
    read(u, *)
    if (nbody > 0) then
      do j = 1, nst
        read(u, '(I5,I8)') ist, n
        if (ist < 1 .or. ist > nst .or. n < 0 .or. n > nbody) goto 900
        cfg%mblist(ist) = n
      end do
    end if
    read(u, *)
    do k = 1, nbody
      do i = 1, ns
        read(u, '(2I5,1PE12.4)') n, isp, diffmu_r
        if (n < 1 .or. n > nbody .or. isp < 1 .or. isp > ns) goto 900
        cfg%effy3b(isp, n) = diffmu_r
      end do
    end do
    ! Gibbs, Lindemann, Troe, SRI step lists and data
    read(u, *)
    read(u, '(I5)') ngibb
    if (ngibb > 0) then
      do j = 1, nst
        read(u, '(I5,I8)') ist, n
        if (ist < 1 .or. ist > nst) goto 900
        cfg%mglist(ist) = n
      end do
    end if
    read(u, *)
    read(u, '(I5)') nlind
    if (nlind < 0 .or. nlind > SG_MAX_FALL) goto 900
    if (nlind > 0) then
      do j = 1, nst
        read(u, '(I5,I8)') ist, n
        if (ist < 1 .or. ist > nst .or. n < 0 .or. n > nlind) goto 900
        cfg%mllist(ist) = n
      end do
    end if
    read(u, *)
    do k = 1, nlind
      read(u, '(I5,4(1PE12.4))') n, rdat(1:4)
      if (n < 1 .or. n > nlind) goto 900
      cfg%rclind(:, n) = rdat(1:4)
    end do
    read(u, *)
    read(u, '(I5)') ntroe
    if (ntroe < 0 .or. ntroe > SG_MAX_FALL) goto 900
    if (ntroe > 0) then
      do j = 1, nst
        read(u, '(I5,I8)') ist, n
        if (ist < 1 .or. ist > nst .or. n < 0 .or. n > ntroe) goto 900
        cfg%mtlist(ist) = n
      end do
    end if
    read(u, *)
    do k = 1, ntroe
      read(u, '(I5,6(1PE12.4))') n, rdat(1:6)
      read(u, '(I5,6(1PE12.4))') n, rdat(7:12)
      if (n < 1 .or. n > ntroe) goto 900
      cfg%rctroe(:, n) = rdat(1:12)
    end do
    read(u, *)
    read(u, '(I5)') nsrif
    if (nsrif < 0 .or. nsrif > SG_MAX_FALL) goto 900
    if (nsrif > 0) then
      do j = 1, nst
        read(u, '(I5,I8)') ist, n
        if (ist < 1 .or. ist > nst .or. n < 0 .or. n > nsrif) goto 900
        cfg%mslist(ist) = n
      end do
    end if
    read(u, *)
    do k = 1, nsrif
      read(u, '(I5,3(1PE12.4))') n, rdat(1:3)
      read(u, '(I5,5(1PE12.4))') n, rdat(4:8)
      if (n < 1 .or. n > nsrif) goto 900
      cfg%rcsrif(:, n) = rdat(1:8)
    end do
    close(u)

    ! ---- derived tables, src/chemin.f:822-983
    do j = 1, nst
      cfg%rparam(1, j) = log(cfg%rparam(1, j))
      cfg%rparam(3, j) = cfg%rparam(3, j) / RGUNIV
      do k = 1, cfg%nsslen(j)
        cfg%diffmw(k, j) = cfg%diffmu(k, j) * cfg%wmolar(cfg%nsspec(k, j) + 1)
      end do
    end do
    do k = 1, nlind
      cfg%rclind(1, k) = log(cfg%rclind(1, k))
      cfg%rclind(3, k) = cfg%rclind(3, k) / RGUNIV
    end do
    do k = 1, ntroe
      cfg%rctroe(1, k) = log(cfg%rctroe(1, k))
      cfg%rctroe(3, k) = cfg%rctroe(3, k) / RGUNIV
      cfg%rctroe(5, k) = -1.0d0 / cfg%rctroe(5, k)
      cfg%rctroe(7, k) = -1.0d0 / cfg%rctroe(7, k)
    end do
    do k = 1, nsrif
      cfg%rcsrif(1, k) = log(cfg%rcsrif(1, k))
      cfg%rcsrif(3, k) = cfg%rcsrif(3, k) / RGUNIV
      cfg%rcsrif(5, k) = -cfg%rcsrif(5, k)
      cfg%rcsrif(6, k) = -1.0d0 / cfg%rcsrif(6, k)
    end do
    if (nbody == 0) cfg%mblist = 0
    giblet = log(prefgb / RGUNIV)
    do i = 1, ns
      cfg%ovwmol(i) = 1.0d0 / cfg%wmolar(i)
      cfg%rgspec(i) = RGUNIV * cfg%ovwmol(i)
      cfg%olewis(i) = 1.0d0 / clewis(i)
      do it = 1, cfg%ntint(i)
        nc = ncofcp(it, i)
        cfg%ncpoly(it, i) = nc - 2
        cfg%ncenth(it, i) = nc - 1
        cfg%ncenpy(it, i) = nc
        do icp = 1, nc
          cfg%amascp(icp, it, i) = amolcp(icp, it, i) * RGUNIV * cfg%ovwmol(i)
        end do
        cfg%amasct(1, it, i) = cfg%amascp(1, it, i) - cfg%rgspec(i)
        cfg%amasch(1, it, i) = cfg%amascp(1, it, i)
        cfg%amolgb(1, it, i) = amolcp(nc, it, i) - amolcp(1, it, i) + giblet
        do icp = 2, nc - 2
          cfg%amasct(icp, it, i) = cfg%amascp(icp, it, i) / real(icp, c_double)
          cfg%amasch(icp, it, i) = cfg%amascp(icp, it, i) / real(icp, c_double)
          cfg%amolgb(icp, it, i) = amolcp(icp, it, i) / real(icp * (icp - 1), c_double)
        end do
        cfg%amasct(nc - 1, it, i) = 0.0d0
        cfg%amasct(nc, it, i) = 0.0d0
        cfg%amasch(nc - 1, it, i) = cfg%amascp(nc - 1, it, i)
        cfg%amasch(nc, it, i) = 0.0d0
        cfg%amolgb(nc - 1, it, i) = amolcp(nc - 1, it, i)
        cfg%amolgb(nc, it, i) = amolcp(1, it, i) - 1.0d0
      end do
    end do
    cfg%alamda = ALAMDC * exp(-RLAMDA * log(TLAMDA))
    cfg%rlamda = RLAMDA
    cfg%prantl = PRANTL
    cfg%clnten = log(1.0d1)                 ! src/indata.f:1071
    ierr = 0
    return
    ! a step, list position, species, third-body or falloff index out of range
900 ierr = 2
    close(u)
  end subroutine read_chem

  ! ---------------------------------------------------------------- diff.dat
  ! Mixture-averaged transport fits (src/diffin.f:95-212) and the symmetric
  ! binary-diffusion fill and Wilke factors (src/diffin.f:280-405).
  subroutine read_diff(path, cfg, ierr)
    character(len=*), intent(in) :: path
    type(sg_config_t), intent(inout) :: cfg
    integer, intent(out) :: ierr
    integer :: u, ns, i, j, k, n, jsp, ncv, ncc, ncd, nct
    character(len=32) :: sym
    real(c_double) :: fornow
    open(newunit=u, file=path, status='old', form='formatted', iostat=ierr)
    if (ierr /= 0) return
    do i = 1, 6
      read(u, *)
    end do
    read(u, '(I5)') ns
    do i = 1, ns
      read(u, '(I5,3X,A)') n, sym
    end do
    read(u, *)
    read(u, *) cfg%pdifgb, cfg%tdifgb
    read(u, *)
    do i = 1, ns
      read(u, *) jsp, ncv
      if (ncv > SG_NTRC) then
        ierr = 1
        close(u)
        return
      end if
      read(u, *) cfg%viscco(1:ncv, i)
      if (i == 1) cfg%ncovis = ncv
    end do
    read(u, *)
    do i = 1, ns
      read(u, *) jsp, ncc
      if (ncc > SG_NTRC) then
        ierr = 1
        close(u)
        return
      end if
      read(u, *) cfg%condco(1:ncc, i)
      if (i == 1) cfg%ncocon = ncc
    end do
    read(u, *)
    do i = 1, ns
      read(u, *) jsp, ncd
      if (ncd > SG_NTRC) then
        ierr = 1
        close(u)
        return
      end if
      do j = 1, i
        read(u, *) k, cfg%diffco(1:ncd, j, i)
      end do
      if (i == 1) cfg%ncodif = ncd
    end do
    ! thermal diffusion ratios: species i lists j < i (src/diffin.f:180-195)
    read(u, *)
    do i = 1, ns
      read(u, *) jsp, nct
      if (nct > SG_NTRC) then
        ierr = 1
        close(u)
        return
      end if
      do j = 1, i - 1
        read(u, *) k, cfg%tdrcco(1:nct, j, i)
      end do
      if (i == 1) cfg%ncotdr = nct
    end do
    close(u)
    do i = 1, ns
      do j = i + 1, ns
        cfg%diffco(:, j, i) = cfg%diffco(:, i, j)
      end do
    end do
    ! antisymmetric fill in the reference's order (src/diffin.f:345-368)
    do i = 1, ns
      cfg%tdrcco(:, i, i) = 0.0d0
      do j = 1, i - 1
        cfg%tdrcco(:, j, i) = -cfg%tdrcco(:, j, i)
      end do
      do j = i + 1, ns
        cfg%tdrcco(:, j, i) = cfg%tdrcco(:, i, j)
      end do
    end do
    do i = 1, ns
      do j = 1, ns
        fornow = cfg%wmolar(i) / cfg%wmolar(j)
        cfg%wilko2(j, i) = 1.0d0 / sqrt(sqrt(fornow))
        fornow = 8.0d0 * (1.0d0 + fornow)
        cfg%wilko1(j, i) = 1.0d0 / sqrt(fornow)
      end do
    end do
    ierr = 0
  end subroutine read_diff

  ! Optically thin radiation data (src/radcin.f:76-126): the radiating species
  ! (0-based indices in sg_config) and their Planck mean absorption
  ! coefficient polynomials.
  subroutine read_radn(path, cfg, ierr)
    character(len=*), intent(in) :: path
    type(sg_config_t), intent(inout) :: cfg
    integer, intent(out) :: ierr
    integer :: u, ns, i, n, jsp, nk
    character(len=32) :: sym
    open(newunit=u, file=path, status='old', form='formatted', iostat=ierr)
    if (ierr /= 0) return
    do i = 1, 6
      read(u, *)
    end do
    read(u, '(I5)') ns
    do i = 1, ns
      read(u, '(I5,3X,A)') n, sym
    end do
    read(u, *)
    read(u, *) cfg%nsprad
    if (cfg%nsprad < 0 .or. cfg%nsprad > SG_MAX_SPEC) then
      ierr = 1
      close(u)
      return
    end if
    do i = 1, cfg%nsprad
      read(u, *) jsp, nk
      if (nk < 1 .or. nk > SG_NKPRAD) then
        ierr = 1
        close(u)
        return
      end if
      read(u, *) cfg%akprad(1:nk, i)
      cfg%nsprid(i) = jsp - 1
      cfg%nkprad(i) = nk
    end do
    close(u)
    ierr = 0
  end subroutine read_radn

  ! Grid, boundary conditions, default state and transport switch from cont.dat.
  subroutine fill_config(ctl, cfg)
    type(control_t), intent(in) :: ctl
    type(sg_config_t), intent(inout) :: cfg
    integer :: i
    cfg%nx = ctl%nx
    cfg%ny = ctl%ny
    cfg%nz = ctl%nz
    cfg%mixavg = merge(1, 0, ctl%nfmavt == 1)
    ! optional physics that is switched on: the molecular-transport options
    ! act only with mixture-averaged transport (src/diffin.f:66,384-407).
    ! sg_create refuses none of them.
    cfg%physics = 0
    if (ctl%nfmavt == 1) then
      if (ctl%nfmixw == 1) cfg%physics = ior(cfg%physics, SG_PHYS_MIXW)
      if (ctl%nfmixp == 1) cfg%physics = ior(cfg%physics, SG_PHYS_MIXP)
      if (ctl%nfmsor == 1) cfg%physics = ior(cfg%physics, SG_PHYS_SORET)
      if (ctl%nfmduf == 1) cfg%physics = ior(cfg%physics, SG_PHYS_DUFOUR)
    end if
    if (ctl%nfradn == 1) cfg%physics = ior(cfg%physics, SG_PHYS_RADN)
    cfg%precision = 0
    cfg%wmltdr = ctl%wmltdr
    cfg%trefrn = ctl%trefrn
    cfg%xgdlen = ctl%xgdlen
    cfg%ygdlen = ctl%ygdlen
    cfg%zgdlen = ctl%zgdlen
    do i = 1, 6
      cfg%bc_type(i) = ctl%bctype(i)
      cfg%bc_ipar(:, i) = ctl%bcipar(:, i)
      cfg%bc_rpar(:, i) = ctl%bcrpar(:, i)
    end do
    cfg%prin = ctl%prin
    cfg%trin = ctl%trin
    cfg%urin = ctl%urin
    cfg%vrin = ctl%vrin
    cfg%wrin = ctl%wrin
    do i = 1, cfg%nspec
      cfg%yrin(i) = ctl%yrin(i)
    end do
  end subroutine fill_config

  ! ---------------------------------------------------------------- dumps
  ! run(nx*ny*nz, 5+N): DRUN, URUN, VRUN, WRUN, ERUN, YRUN(:,1..N).
  subroutine read_dump(path, formatted, nx, ny, nz, ns, run, etime, tstep, errold, errldr, ierr)
    character(len=*), intent(in) :: path
    logical, intent(in) :: formatted
    integer, intent(in) :: nx, ny, nz, ns
    real(c_double), intent(out) :: run(nx * ny * nz, 5 + ns)
    real(c_double), intent(out) :: etime, tstep, errold, errldr
    integer, intent(out) :: ierr
    integer :: u, n1, n2, n3, n4, p
    if (formatted) then
      open(newunit=u, file=path, status='old', form='formatted', iostat=ierr)
      if (ierr /= 0) return
      read(u, *) n1, n2, n3, n4
      do p = 1, nx * ny * nz
        read(u, *) run(p, :)
      end do
      read(u, *) etime, tstep, errold, errldr
    else
      open(newunit=u, file=path, status='old', form='unformatted', iostat=ierr)
      if (ierr /= 0) return
      read(u) n1, n2, n3, n4, run, etime, tstep, errold, errldr
    end if
    close(u)
    ierr = 0
    if (n1 /= nx .or. n2 /= ny .or. n3 /= nz .or. n4 /= ns) ierr = 2
  end subroutine read_dump

  subroutine write_dump(path, formatted, nx, ny, nz, ns, run, etime, tstep, errold, errldr, ierr)
    character(len=*), intent(in) :: path
    logical, intent(in) :: formatted
    integer, intent(in) :: nx, ny, nz, ns
    real(c_double), intent(in) :: run(nx * ny * nz, 5 + ns)
    real(c_double), intent(in) :: etime, tstep, errold, errldr
    integer, intent(out) :: ierr
    integer :: u, p
    if (formatted) then
      open(newunit=u, file=path, status='replace', form='formatted', iostat=ierr)
      if (ierr /= 0) return
      write(u, *) nx, ny, nz, ns
      do p = 1, nx * ny * nz
        write(u, *) run(p, :)
      end do
      write(u, *) etime, tstep, errold, errldr
    else
      open(newunit=u, file=path, status='replace', form='unformatted', iostat=ierr)
      if (ierr /= 0) return
      write(u) nx, ny, nz, ns, run, etime, tstep, errold, errldr
    end if
    close(u)
    ierr = 0
  end subroutine write_dump

  ! Grid and species count of a dump, from its header (both forms).
  subroutine dump_header(path, formatted, nx, ny, nz, ns, ierr)
    character(len=*), intent(in) :: path
    logical, intent(in) :: formatted
    integer, intent(out) :: nx, ny, nz, ns, ierr
    integer :: u
    if (formatted) then
      open(newunit=u, file=path, status='old', form='formatted', iostat=ierr)
      if (ierr /= 0) return
      read(u, *, iostat=ierr) nx, ny, nz, ns
    else
      open(newunit=u, file=path, status='old', form='unformatted', iostat=ierr)
      if (ierr /= 0) return
      read(u, iostat=ierr) nx, ny, nz, ns
    end if
    close(u)
  end subroutine dump_header

  ! ------------------------------------------------------------- run output
  ! output/rept.res: created at start-up (src/indata.f:216-230 writes a
  ! header); the loop appends to it (src/output.f:81-99, 343-352).
  subroutine report_open(path, ctl, ns, tstep)
    character(len=*), intent(in) :: path
    type(control_t), intent(in) :: ctl
    integer, intent(in) :: ns
    real(c_double), intent(in) :: tstep
    integer :: u
    character(len=72) :: stars
    stars = repeat('*', 72)
    open(newunit=u, file=path, status='replace', form='formatted')
    write(u, '(a)') stars
    write(u, '(a)') '*  SENGA+ right-hand side on one MI355X (senga_gpu driver)'
    write(u, '(a)') stars
    write(u, *)
    write(u, *) '==========='
    write(u, *) 'REPORT FILE'
    write(u, *) '==========='
    write(u, *)
    write(u, '(a, 3i6)') ' Global grid (nx,ny,nz): ', ctl%nx, ctl%ny, ctl%nz
    write(u, '(a, 3(1pe12.4))') ' Domain size (m): ', ctl%xgdlen, ctl%ygdlen, ctl%zgdlen
    write(u, '(a, i4)') ' Species: ', ns
    write(u, '(a, 1pe12.4, a, i7, a, i7, a, i2)') ' Time step: ', tstep, '; first step', ctl%ntime1, &
        '; steps', ctl%ntime, '; step switch', ctl%nstpsw
    write(u, '(a, 6i4)') ' Boundary types: ', ctl%bctype
    write(u, *)
    close(u)
  end subroutine report_open

  ! every NTREPT steps: format 9100 and 9110 of src/output.f:440-441
  subroutine report_step(path, itime, etime, tstep)
    character(len=*), intent(in) :: path
    integer, intent(in) :: itime
    real(c_double), intent(in) :: etime, tstep
    integer :: u
    open(newunit=u, file=path, status='old', position='append', form='formatted')
    write(u, '(''Time step number: '', i7)') itime
    write(u, '(''Elapsed time: '', 1pe12.4, '';'', 2x, ''next time step:'', 1pe12.4)') etime, tstep
    close(u)
  end subroutine report_step

  ! after each dump: format 9120 of src/output.f:442
  subroutine report_dump(path, fname)
    character(len=*), intent(in) :: path, fname
    integer :: u
    open(newunit=u, file=path, status='old', position='append', form='formatted')
    write(u, '(''Dump completed: '', a)') fname
    close(u)
  end subroutine report_dump

  ! output/stat.res: created empty at start-up (src/indata.f:1024-1026); the
  ! step number appended every NTSTAT steps (src/output.f:397-425)
  subroutine stat_open(path)
    character(len=*), intent(in) :: path
    integer :: u
    open(newunit=u, file=path, status='replace', form='formatted')
    close(u)
  end subroutine stat_open

  subroutine stat_step(path, itime)
    character(len=*), intent(in) :: path
    integer, intent(in) :: itime
    integer :: u
    open(newunit=u, file=path, status='old', position='append', form='formatted')
    write(u, '(''Time step number: '', i7)') itime
    close(u)
  end subroutine stat_step

  ! output/outIIIIIPPPP.res on every dump step (src/output.f:241-265): one
  ! unformatted record of rho, u, v, w, E per unit mass (ERUN/DRUN), T, p,
  ! Y_k, the reaction rates of the last right-hand side and ETIME.
  subroutine write_out_res(path, nx, ny, nz, ns, run, t, p, rate, etime, ierr)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nx, ny, nz, ns
    real(c_double), intent(in) :: run(nx * ny * nz, 5 + ns), t(nx * ny * nz), p(nx * ny * nz)
    real(c_double), intent(in) :: rate(nx * ny * nz, ns)
    real(c_double), intent(in) :: etime
    integer, intent(out) :: ierr
    integer :: u, k
    open(newunit=u, file=path, status='replace', form='unformatted', iostat=ierr)
    if (ierr /= 0) return
    write(u) run(:, 1), run(:, 2) / run(:, 1), run(:, 3) / run(:, 1), run(:, 4) / run(:, 1), &
             run(:, 5) / run(:, 1), t, p, ((run(:, 5 + k) / run(:, 1)), k = 1, ns), rate, etime
    close(u)
  end subroutine write_out_res

end module senga_io
