! senga_gpu_mod.f90 -- ISO_C_BINDING interface to libsenga_gpu (include/senga_gpu.h).
!
! A SENGA+ time loop (src/senga2.f:96-194) drives the GPU path through these
! interfaces: sg_substage(h, irk) in place of
!   BOUNDT; PARFER; RHSCAL; RHSVEL; BOUNDS; LINCOM/FINCOM   (src/senga2.f:113-159)
! and sg_end_step(h) in place of ETIME = ETIME + TSTEP; BOUNTT (src/senga2.f:163-171).
!
! sg_config_t mirrors the C struct sg_config field for field.  C arrays
! a[i][j][k] appear here as a(k, j, i): both are the same bytes.
module senga_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: SG_MAX_SPEC = 18, SG_MAX_STEP = 128, SG_MAX_SSL = 10
  integer, parameter, public :: SG_MAX_BODY = 16, SG_MAX_FALL = 32
  integer, parameter, public :: SG_NCOF = 7, SG_NTINT = 2, SG_NTRC = 4, SG_NKPRAD = 6
  integer, parameter, public :: SG_OK = 0, SG_EINVAL = 1, SG_ENOMEM = 2, SG_EDEVICE = 3
  integer, parameter, public :: SG_ETEMPER = 4, SG_ETEMPIN = 5, SG_ESTATE = 6
  integer, parameter, public :: SG_BC_PERIODIC = 1, SG_BC_INFLOW3 = 13, SG_BC_OUTFLOW1 = 21
  ! sg_config%physics bits: optional physics switched on in cont.dat
  integer, parameter, public :: SG_PHYS_MIXW = 1, SG_PHYS_MIXP = 2, SG_PHYS_SORET = 4
  integer, parameter, public :: SG_PHYS_DUFOUR = 8, SG_PHYS_RADN = 16

  type, bind(C), public :: sg_config_t
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: nspec, nstep
    integer(c_int32_t) :: mixavg
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: physics
    real(c_double) :: xgdlen, ygdlen, zgdlen
    integer(c_int32_t) :: bc_type(6)
    integer(c_int32_t) :: bc_ipar(4, 6)
    real(c_double) :: bc_rpar(4, 6)
    real(c_double) :: prin, trin, urin, vrin, wrin
    real(c_double) :: yrin(SG_MAX_SPEC)
    real(c_double) :: alamda, rlamda, prantl
    real(c_double) :: olewis(SG_MAX_SPEC)
    real(c_double) :: wmolar(SG_MAX_SPEC), ovwmol(SG_MAX_SPEC), rgspec(SG_MAX_SPEC)
    integer(c_int32_t) :: ntint(SG_MAX_SPEC)
    integer(c_int32_t) :: ncpoly(SG_NTINT, SG_MAX_SPEC)
    integer(c_int32_t) :: ncenth(SG_NTINT, SG_MAX_SPEC)
    integer(c_int32_t) :: ncenpy(SG_NTINT, SG_MAX_SPEC)
    real(c_double) :: tinthi(SG_NTINT, SG_MAX_SPEC)
    real(c_double) :: amascp(SG_NCOF, SG_NTINT, SG_MAX_SPEC)
    real(c_double) :: amasct(SG_NCOF, SG_NTINT, SG_MAX_SPEC)
    real(c_double) :: amasch(SG_NCOF, SG_NTINT, SG_MAX_SPEC)
    real(c_double) :: amolgb(SG_NCOF, SG_NTINT, SG_MAX_SPEC)
    real(c_double) :: rparam(3, SG_MAX_STEP)
    integer(c_int32_t) :: nsslen(SG_MAX_STEP)
    integer(c_int32_t) :: nsspec(SG_MAX_SSL, SG_MAX_STEP)
    real(c_double) :: diffmu(SG_MAX_SSL, SG_MAX_STEP)
    real(c_double) :: diffmw(SG_MAX_SSL, SG_MAX_STEP)
    integer(c_int32_t) :: nrslen(SG_MAX_STEP)
    integer(c_int32_t) :: nrspec(SG_MAX_SSL, SG_MAX_STEP)
    integer(c_int32_t) :: npslen(SG_MAX_STEP)
    integer(c_int32_t) :: npspec(SG_MAX_SSL, SG_MAX_STEP)
    integer(c_int32_t) :: nrclen(SG_MAX_STEP)
    integer(c_int32_t) :: nrcpec(SG_MAX_SSL, SG_MAX_STEP)
    real(c_double) :: crspec(SG_MAX_SSL, SG_MAX_STEP)
    integer(c_int32_t) :: npclen(SG_MAX_STEP)
    integer(c_int32_t) :: npcpec(SG_MAX_SSL, SG_MAX_STEP)
    real(c_double) :: cpspec(SG_MAX_SSL, SG_MAX_STEP)
    integer(c_int32_t) :: mblist(SG_MAX_STEP)
    integer(c_int32_t) :: mglist(SG_MAX_STEP)
    integer(c_int32_t) :: mllist(SG_MAX_STEP)
    integer(c_int32_t) :: mtlist(SG_MAX_STEP)
    integer(c_int32_t) :: mslist(SG_MAX_STEP)
    real(c_double) :: effy3b(SG_MAX_SPEC, SG_MAX_BODY)
    real(c_double) :: rclind(4, SG_MAX_FALL)
    real(c_double) :: rctroe(12, SG_MAX_FALL)
    real(c_double) :: rcsrif(8, SG_MAX_FALL)
    real(c_double) :: clnten
    real(c_double) :: pdifgb, tdifgb
    integer(c_int32_t) :: ncovis, ncocon, ncodif, reserved1
    real(c_double) :: viscco(SG_NTRC, SG_MAX_SPEC)
    real(c_double) :: condco(SG_NTRC, SG_MAX_SPEC)
    real(c_double) :: diffco(SG_NTRC, SG_MAX_SPEC, SG_MAX_SPEC)
    real(c_double) :: wilko1(SG_MAX_SPEC, SG_MAX_SPEC)
    real(c_double) :: wilko2(SG_MAX_SPEC, SG_MAX_SPEC)
    integer(c_int32_t) :: ncotdr, reserved2
    real(c_double) :: wmltdr
    real(c_double) :: tdrcco(SG_NTRC, SG_MAX_SPEC, SG_MAX_SPEC)
    integer(c_int32_t) :: nsprad, reserved3
    integer(c_int32_t) :: nsprid(SG_MAX_SPEC)
    integer(c_int32_t) :: nkprad(SG_MAX_SPEC)
    real(c_double) :: trefrn
    real(c_double) :: akprad(SG_NKPRAD, SG_MAX_SPEC)
  end type sg_config_t

  public :: sg_abi_version, sg_config_size, sg_create, sg_destroy
  public :: sg_upload_state, sg_download_state, sg_set_timestep
  public :: sg_substage, sg_end_step, sg_step, sg_rhs, sg_derivatives
  public :: sg_sync, sg_last_error, sg_memory, sg_set_adaptive, sg_get_timestep
  public :: sg_mechanism_compiled

  interface
    integer(c_int) function sg_abi_version() bind(C, name="sg_abi_version")
      import :: c_int
    end function
    integer(c_size_t) function sg_config_size() bind(C, name="sg_config_size")
      import :: c_size_t
    end function
    ! int sg_create(const sg_config *cfg, sg_handle **out)
    integer(c_int) function sg_create(cfg, h) bind(C, name="sg_create")
      import :: c_int, c_ptr, sg_config_t
      type(sg_config_t), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function sg_destroy(h) bind(C, name="sg_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! run(nx,ny,nz,5+N): DRUN, URUN, VRUN, WRUN, ERUN, YRUN (src/com_senga2.h:576-583)
    integer(c_int) function sg_upload_state(h, run, etime) bind(C, name="sg_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: run(*)
      real(c_double), value :: etime
    end function
    ! Optional outputs: pass c_null_ptr through the c_ptr form to skip one.
    integer(c_int) function sg_download_state(h, run, temp, pres, rate) bind(C, name="sg_download_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, run, temp, pres, rate
    end function
    integer(c_int) function sg_set_timestep(h, tstep) bind(C, name="sg_set_timestep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: tstep
    end function
    integer(c_int) function sg_substage(h, irk) bind(C, name="sg_substage")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: irk
    end function
    integer(c_int) function sg_end_step(h) bind(C, name="sg_end_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function sg_step(h, nsteps) bind(C, name="sg_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function sg_rhs(h, irk, rhs) bind(C, name="sg_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: irk
      real(c_double), intent(out) :: rhs(*)
    end function
    integer(c_int) function sg_derivatives(h, f, out) bind(C, name="sg_derivatives")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function sg_sync(h) bind(C, name="sg_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function sg_last_error(h, code, ijk, msg, n) bind(C, name="sg_last_error")
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: code
      integer(c_int), intent(out) :: ijk(3)
      character(kind=c_char), intent(out) :: msg(*)
      integer(c_size_t), value :: n
    end function
    ! cont.dat step switch 1: ADAPTT after every step (src/adaptt.f)
    integer(c_int) function sg_set_adaptive(h, on, errold, errldr) bind(C, name="sg_set_adaptive")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: on
      real(c_double), value :: errold, errldr
    end function
    integer(c_int) function sg_get_timestep(h, tstep, errold, errldr) bind(C, name="sg_get_timestep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: tstep, errold, errldr
    end function
    integer(c_int) function sg_memory(h, bytes) bind(C, name="sg_memory")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: h
      integer(c_size_t), intent(out) :: bytes
    end function
    integer(c_int) function sg_mechanism_compiled(h, compiled) bind(C, name="sg_mechanism_compiled")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: compiled
    end function
  end interface
end module senga_gpu
