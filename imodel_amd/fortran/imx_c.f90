!=======================================================================
! imx_c: ISO_C_BINDING interface to libimx (include/imx_swm.h).
! Every libimx call returns 0 or a negative code; imx_check turns a failure
! into `print*, message; stop`, the reference's error convention
! (e.g. src/swm_operators.f90:1604-1606).
!=======================================================================
module imx_c
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: IMX_PREC_FP64_REF = 0, IMX_PREC_FP64 = 1, IMX_PREC_FP32 = 2
  integer(c_int), parameter :: IMX_ENOTSUP = -6
  integer(c_int32_t), parameter :: IMX_FIELD_H_ED = 0, IMX_FIELD_ETA = 2, IMX_FIELD_Q_TR = 4, IMX_FIELD_Q_ED = 5, &
                                   IMX_FIELD_DIVUH = 6, IMX_FIELD_KE_HX = 7, IMX_FIELD_DIVU = 11
  integer(c_int32_t), parameter :: IMX_POS_EQS = 0, IMX_POS_POL = 1
  integer, parameter :: IMX_NDIAG = 8
  integer, parameter :: IMX_LTE_NROW = 18, IMX_LTE_NMIM = 7

  type, bind(C) :: imx_mesh_desc
    integer(c_int32_t) :: nv, ne, nt, glevel
    type(c_ptr) :: v_nnb, v_nb, v_ed, v_tr, hx_ttgout, hx_tnrccw, v_p, hx_areag
    type(c_ptr) :: ed_v, ed_sh, ed_leng, edhx_leng, ed_c, edhx_c
    type(c_ptr) :: tr_v, tr_ed, tr_tg, tr_areag, tr_c, tr_clat
    type(c_ptr) :: tr_trhx_area, hx_hxtr_area, hx_trskw
  end type imx_mesh_desc

  type, bind(C) :: imx_swm_opts
    integer(c_int32_t) :: precision, testcase, use_sfc, use_graph
    real(c_double) :: erad, grav, omega, dt, h_ref
  end type imx_swm_opts

  ! imx_scheme: APVM, diffusion, hyperdiffusion (IMX_COEF_*: const, align, diam)
  integer(c_int32_t), parameter :: IMX_COEF_CONST = 0, IMX_COEF_ALIGN = 1, IMX_COEF_DIAM = 2
  integer(c_int32_t), parameter :: IMX_WRAP_TRSK10 = 0, IMX_WRAP_PXT16 = 1, IMX_WRAP_GASS18 = 2, &
                                   IMX_WRAP_NONE = 3
  ! the individual methods of wrapper none: scalar interpolation, Coriolis,
  ! kinetic energy
  integer(c_int32_t), parameter :: IMX_SINT_TRSK = 0, IMX_SINT_BARY = 1, IMX_SINT_GASS = 2
  integer(c_int32_t), parameter :: IMX_COR_TRSK = 0, IMX_COR_PERED = 1, IMX_COR_GASS = 2, IMX_COR_DTRED = 3
  integer(c_int32_t), parameter :: IMX_REC_TRSK = 0, IMX_REC_PERHX = 1, IMX_REC_GASS = 2, IMX_REC_MELV = 3
  type, bind(C) :: imx_scheme
    integer(c_int32_t) :: stag, wrap, apvm
    real(c_double) :: pvspar, k2
    integer(c_int32_t) :: diffus
    real(c_double) :: k4
    integer(c_int32_t) :: hyperdiffus
    integer(c_int32_t) :: sinterp = 0, coriolis = 0, recon = 0
    real(c_double) :: gasscoef = 1._c_double
    integer(c_int32_t) :: nopv = 0
  end type imx_scheme

  interface
    function imx_last_error() bind(C, name='imx_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function imx_mesh_build_icos(glevel, pos, out) bind(C, name='imx_mesh_build_icos') result(rc)
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: glevel, pos
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function imx_mesh_build_read(path, out) bind(C, name='imx_mesh_build_read') result(rc)
      import :: c_int, c_char, c_ptr
      character(kind=c_char) :: path(*)
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function imx_mesh_gridprop(m) bind(C, name='imx_mesh_gridprop') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int) :: rc
    end function
    function imx_mesh_desc_of(m, d) bind(C, name='imx_mesh_desc_of') result(rc)
      import :: c_int, c_ptr, imx_mesh_desc
      type(c_ptr), value :: m
      type(imx_mesh_desc) :: d
      integer(c_int) :: rc
    end function
    ! meshload / meshwrite grid files (<dir>/<name>_*.dat), C strings
    function imx_mesh_load_ref(dir, name, out) bind(C, name='imx_mesh_load_ref') result(rc)
      import :: c_int, c_ptr, c_char
      character(kind=c_char) :: dir(*), name(*)
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function imx_mesh_write_ref(m, dir, name) bind(C, name='imx_mesh_write_ref') result(rc)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: m
      character(kind=c_char) :: dir(*), name(*)
      integer(c_int) :: rc
    end function
    function imx_remap_ll(d, pos, f, nlon, nlat, out) bind(C, name='imx_remap_ll') result(rc)
      import :: c_int, c_int32_t, c_double, c_float, imx_mesh_desc
      type(imx_mesh_desc) :: d
      integer(c_int32_t), value :: pos, nlon, nlat
      real(c_double) :: f(*)
      real(c_float) :: out(*)
      integer(c_int) :: rc
    end function
    function imx_swm_jacobian(s, d, h, u, pert, cap, row, col, val, nnz, ncolors) &
        bind(C, name='imx_swm_jacobian') result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, imx_mesh_desc
      type(c_ptr), value :: s
      type(imx_mesh_desc) :: d
      real(c_double) :: h(*), u(*)
      real(c_double), value :: pert
      integer(c_int64_t), value :: cap
      integer(c_int32_t) :: row(*), col(*)
      real(c_double) :: val(*)
      integer(c_int64_t) :: nnz
      integer(c_int32_t) :: ncolors
      integer(c_int) :: rc
    end function
    function imx_make_dir(path) bind(C, name='imx_make_dir') result(rc)
      import :: c_int, c_char
      character(kind=c_char) :: path(*)
      integer(c_int) :: rc
    end function
    subroutine imx_mesh_destroy(m) bind(C, name='imx_mesh_destroy')
      import :: c_ptr
      type(c_ptr), value :: m
    end subroutine
    function imx_ic_tc2_pv(m, erad, q) bind(C, name='imx_ic_tc2_pv') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), value :: erad
      real(c_double) :: q(*)
      integer(c_int) :: rc
    end function
    function imx_swm_get_field(s, which, out) bind(C, name='imx_swm_get_field') result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: s
      integer(c_int32_t), value :: which
      real(c_double) :: out(*)
      integer(c_int) :: rc
    end function
    function imx_ic_tc2_hc(m, erad, h, u) bind(C, name='imx_ic_tc2_hc') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), value :: erad
      real(c_double) :: h(*), u(*)
      integer(c_int) :: rc
    end function
    function imx_ic_tc2(m, erad, h, u) bind(C, name='imx_ic_tc2') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: m
      real(c_double), value :: erad
      real(c_double) :: h(*), u(*)
      integer(c_int) :: rc
    end function
    function imx_ic_spectral(d, seed, lmax, h0, hfrac, urms, erad, h, u) &
        bind(C, name='imx_ic_spectral') result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double, imx_mesh_desc
      type(imx_mesh_desc) :: d
      integer(c_int64_t), value :: seed
      integer(c_int32_t), value :: lmax
      real(c_double), value :: h0, hfrac, urms, erad
      real(c_double) :: h(*), u(*)
      integer(c_int) :: rc
    end function
    function imx_ic_testcase(m, tc, stag, erad, hollgw, h, u, bt) bind(C, name='imx_ic_testcase') result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: m
      integer(c_int32_t), value :: tc, stag
      real(c_double), value :: erad, hollgw
      real(c_double) :: h(*), u(*), bt(*)
      integer(c_int) :: rc
    end function
    function imx_ic_testcase_coriolis(tc, fsphere, fcte) bind(C, name='imx_ic_testcase_coriolis') result(rc)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: tc
      integer(c_int32_t) :: fsphere
      real(c_double) :: fcte
      integer(c_int) :: rc
    end function
    function imx_swm_set_coriolis(s, fsphere, fcte) bind(C, name='imx_swm_set_coriolis') result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: s
      integer(c_int32_t), value :: fsphere
      real(c_double), value :: fcte
      integer(c_int) :: rc
    end function
    function imx_swm_set_topography(s, bt) bind(C, name='imx_swm_set_topography') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double) :: bt(*)
      integer(c_int) :: rc
    end function
    subroutine imx_swm_default_opts(o) bind(C, name='imx_swm_default_opts')
      import :: imx_swm_opts
      type(imx_swm_opts) :: o
    end subroutine
    function imx_swm_create(d, o, out) bind(C, name='imx_swm_create') result(rc)
      import :: c_int, c_ptr, imx_mesh_desc, imx_swm_opts
      type(imx_mesh_desc) :: d
      type(imx_swm_opts) :: o
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function imx_swm_set_scheme(s, d, sc) bind(C, name='imx_swm_set_scheme') result(rc)
      import :: c_int, c_ptr, imx_mesh_desc, imx_scheme
      type(c_ptr), value :: s
      type(imx_mesh_desc) :: d
      type(imx_scheme) :: sc
      integer(c_int) :: rc
    end function
    function imx_swm_set_state(s, h, u) bind(C, name='imx_swm_set_state') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double) :: h(*), u(*)
      integer(c_int) :: rc
    end function
    function imx_swm_get_state(s, h, u) bind(C, name='imx_swm_get_state') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double) :: h(*), u(*)
      integer(c_int) :: rc
    end function
    function imx_swm_tendency(s, h, u, masseq, momeq) bind(C, name='imx_swm_tendency') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double) :: h(*), u(*), masseq(*), momeq(*)
      integer(c_int) :: rc
    end function
    function imx_swm_step(s, nsteps) bind(C, name='imx_swm_step') result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: s
      integer(c_int32_t), value :: nsteps
      integer(c_int) :: rc
    end function
    function imx_swm_lte(s, m, err, mim) bind(C, name='imx_swm_lte') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s, m
      real(c_double) :: err(*), mim(*)
      integer(c_int) :: rc
    end function
    function imx_swm_lte_hol(s, m, hollgw, err, mim) bind(C, name='imx_swm_lte_hol') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s, m
      real(c_double), value :: hollgw
      real(c_double) :: err(*), mim(*)
      integer(c_int) :: rc
    end function
    function imx_swm_diagnostics(s, out) bind(C, name='imx_swm_diagnostics') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double) :: out(*)
      integer(c_int) :: rc
    end function
    subroutine imx_swm_destroy(s) bind(C, name='imx_swm_destroy')
      import :: c_ptr
      type(c_ptr), value :: s
    end subroutine
  end interface

contains

  subroutine imx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      print*, "ERROR in ", what, ": ", trim(imx_error_message())
      stop 1
    end if
  end subroutine imx_check

  function imx_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ''
    p = imx_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function imx_error_message
end module imx_c
