!=======================================================================
! swm: the shallow-water module of the imodel driver, on the GPU, with the
! reference's call surface (src/swm.f90):
!   swm_tests(meshtmp)                        :102-380
!   swm_horiz_loc_trunc_er(meshtmp)           :382-734
!   initialize_gridprop()                     :979-1041
!   initialize_fields()                       :1043-2414
!   tendency(h, u, masseq, momeq)             :2416-2598
!   ode_rk4(t, h, u, h_new, u_new, dt)        :2600-2683
!   calc_energies(P, K, T, A)                 :2686-2752
!   sumf_areas(f), sumfsq_areas(f)            :3394-3490
!   error_calc(f, fref, fdif, emr, e2, em)    :3492-3543
! with the types of module datastruct and the module globals of swm_data
! (mesh, h, u, bt, masseq, momeq, ...), so a caller written against the
! reference (e.g. moist_swm's `use swm, only: initialize_gridprop,
! sumf_areas, calc_energies, tendency, error_calc`, src/moist_swm.f90:61-66)
! compiles against it. The state lives in HBM behind one libimx handle,
! opened by initialize_gridprop.
!=======================================================================
module swm
  use iso_c_binding
  use imx_c
  use datastruct
  use swm_data
  implicit none
  private
  public :: swm_tests, swm_horiz_loc_trunc_er, initialize_gridprop, initialize_fields, tendency, ode_rk4, &
            calc_energies, sumf_areas, sumfsq_areas, error_calc
  real(r8), parameter :: erad = 6.37122e6_r8   ! constants.f90:55
  real(r8), parameter :: grav = 9.80616_r8     ! constants.f90:61
  real(r8), parameter :: omega = 7.292e-5_r8   ! constants.f90:67
  real(r8), parameter :: rad2deg = 57.295779513082323_r8

  real(r8), parameter :: sec2day = 1._r8/86400._r8
  type(imx_mesh_desc) :: mdesc
  ! an FP64 reference-order handle on the same mesh and scheme, for the
  ! reference's intermediate fields (div u, K, h_ed) of a state evolved on a
  ! fused handle
  type(c_ptr) :: diag_handle = c_null_ptr
  ! analytic reference solution (RefSolAnal, src/swm_data.f90:685)
  logical :: refsol = .false.
  integer :: blowup = 0, plotsteps = 1
  real(r8), allocatable :: q_exact(:), qtr(:)
  character(len=256) :: swmname = '', meshname = ''
  real(r8) :: cfl = 0._r8, meanvdist = 0._r8

contains

  subroutine swm_tests(meshtmp)
    type(grid_structure), intent(in) :: meshtmp
    real(c_double) :: d0(IMX_NDIAG), d(IMX_NDIAG)
    real(r8), pointer :: leng(:)
    real(r8) :: maxvel, minvdist, time
    integer :: k, kn

    mesh = meshtmp
    call swmpars(usetime=.true.)
    call allocate_globalswmvars()
    call initialize_gridprop()
    call initialize_fields()
    maxvel = maxval(abs(u%f))
    call c_f_pointer(mdesc%ed_leng, leng, [mdesc%ne])
    minvdist = minval(leng)
    meanvdist = sum(leng)/mdesc%ne
    cfl = abs(maxvel*dt/(minvdist*erad))   ! swm.f90:2383
    print*, "CFL:", cfl

    call make_names(mesh%glevel, .true.)

    if (normal_modes == 1) then
      call swm_normal_mode_analysis()
      call close_handles()
      return
    end if
    call imx_check(imx_swm_diagnostics(handle, d0), 'imx_swm_diagnostics')
    call perot_kenergy(d0, .true.)
    ! RefSolAnal (src/swm_data.f90:685): the steady cases compare with the
    ! initial state; TC2 also has an exact PV
    select case (testcase)
    case (1, 2, 22, 32, 33, 34, 35, 40, 41, 42)
      refsol = .true.
    case default
      refsol = .false.
    end select
    if (refsol) then
      allocate(q_exact(mdesc%nt), qtr(mdesc%nt))
      q_exact = 0._r8
      if (testcase == 2) call imx_check(imx_ic_tc2_pv(mesh%lib, erad, q_exact), 'imx_ic_tc2_pv')
    end if
    ! swm.f90:226-231: both files start with the absolute initial values
    call write_evol_error(0, 0._r8, d0(1), 0._r8, 0._r8, 0._r8, 0._r8, 0._r8, 0._r8, &
                          d0(2), d0(3), d0(4), d0(5), 0._r8, 0._r8, 0._r8, 0._r8)
    call write_evol(0, 0._r8, d0(1), d0(2), d0(3), d0(4), d0(5), 0._r8)
    plotsteps = 1
    if (nplots > 0) plotsteps = max(1, ntime/nplots)   ! swm_data.f90:589-597
    call plotfields(0, 0._r8)

    ! time loop of swm_tests (swm.f90:238-370). With the analytic solution
    ! errors are written every nprints steps and at the end
    ! (write_evol_error, write_errors); otherwise the diagnostics at k<=2,
    ! k==ntime and every nprints steps (write_evol). The steps in between
    ! run on the GPU without leaving it, so the reference's per-step
    ! blow-up test is made at the written steps only.
    k = 0
    blowup = 0
    do while (k < ntime)
      kn = min(ntime, (k/nprints + 1)*nprints)
      if (k < 2 .and. .not. refsol) kn = k + 1
      if (nplots > 0) kn = min(kn, (k/plotsteps + 1)*plotsteps)
      call imx_check(imx_swm_step(handle, int(kn - k, c_int32_t)), 'imx_swm_step')
      k = kn
      time = real(k, r8)*dt
      call imx_check(imx_swm_diagnostics(handle, d), 'imx_swm_diagnostics')
      if (d(8) /= 0._r8) then
        print*, "Stopping due to NaN/Inf in the state at step", k
        exit
      end if
      if (refsol) then
        if (mod(k, nprints) == 0 .or. k == ntime) then
          if (.not. write_ref_errors(k, time, d, d0)) exit
        end if
      else if (k <= 2 .or. k == ntime .or. mod(k, nprints) == 0) then
        call perot_kenergy(d, .false.)
        call write_evol(k, time, (d(1) - d0(1))/d0(1), (d(2) - d0(2))/d0(2), (d(3) - d0(3))/d0(3), &
                        (d(4) - d0(4))/d0(4), (d(5) - d0(5))/d0(5), d(7))
      end if
      call plotfields(k, time)
    end do

    call imx_check(imx_swm_get_state(handle, h%f, u%f), 'imx_swm_get_state')
    if (dump_state == 1) call dump_final_state()
    call close_handles()
  end subroutine swm_tests

  ! initialize_gridprop (src/swm.f90:979-1041): the kite areas and TRiSK
  ! weights of the mesh (computed once on the host in FP64), and the device
  ! handle the fields are evolved on, with the scheme of par/swm.par
  subroutine initialize_gridprop()
    call imx_check(imx_mesh_gridprop(mesh%lib), 'initialize_gridprop')
    call imx_check(imx_mesh_desc_of(mesh%lib, mdesc), 'initialize_gridprop')
    call close_handles()
    call open_handle(handle, precision_code())
  end subroutine initialize_gridprop

  ! initialize_fields (src/swm.f90:1043-2414): h, u and bt of the test case
  ! (imx_ic_testcase; 100: the seeded spectral state), loaded on the handle
  subroutine initialize_fields()
    integer(c_int32_t) :: sg
    if (testcase == 100) then
      ! seeded spectral state (new test case 100): h' 5% of H0, 20 m/s rms wind
      call imx_check(imx_ic_spectral(mdesc, int(seed, c_int64_t), 16_c_int32_t, 2.94e4_r8/grav, &
                                     0.05_r8, 20._r8, erad, h%f, u%f), 'initialize_fields')
      bt%f = 0._r8
    else
      sg = merge(1_c_int32_t, 0_c_int32_t, trim(stag) == 'HC')
      call imx_check(imx_ic_testcase(mesh%lib, int(testcase, c_int32_t), sg, erad, hollgw, h%f, u%f, bt%f), &
                     'initialize_fields')
    end if
    call imx_check(imx_swm_set_topography(handle, bt%f), 'initialize_fields')
    call imx_check(imx_swm_set_state(handle, h%f, u%f), 'initialize_fields')
    h_exact%f = h%f
    u_exact%f = u%f
  end subroutine initialize_fields

  ! subroutine tendency(h, u, masseq, momeq)  (src/swm.f90:2416-2598)
  subroutine tendency(h, u, masseq, momeq)
    type(scalar_field), intent(in) :: h, u
    real(r8), intent(inout) :: masseq(:), momeq(:)
    call imx_check(imx_swm_tendency(handle, h%f, u%f, masseq, momeq), 'tendency')
  end subroutine tendency

  ! subroutine ode_rk4(t, h, u, h_new, u_new, dt)  (src/swm.f90:2600-2683):
  ! one RK4 step of the handle's time step from (h, u); t and dt as the
  ! reference takes them (the step length is the handle's, set from dt)
  subroutine ode_rk4(t, h, u, h_new, u_new, dt)
    real(r8) :: t
    type(scalar_field), intent(in) :: h, u
    type(scalar_field) :: h_new, u_new
    real(r8) :: dt
    if (abs(dt - swm_par_dt()) > 0._r8) then
      print*, "ode_rk4: dt", dt, " differs from the run's time step", swm_par_dt()
      stop 1
    end if
    u_new = u
    h_new = h
    call imx_check(imx_swm_set_state(handle, h%f, u%f), 'ode_rk4')
    call imx_check(imx_swm_step(handle, 1_c_int32_t), 'ode_rk4')
    call imx_check(imx_swm_get_state(handle, h_new%f, u_new%f), 'ode_rk4')
    if (t < 0._r8) print*, "ode_rk4: negative time", t
  end subroutine ode_rk4

  real(r8) function swm_par_dt() result(d)
    d = dt
  end function swm_par_dt

  ! sumf_areas (src/swm.f90:3394-3440): area-weighted mean; on edges the
  ! reference divides the sum of cell-area weights by the edge areas, and so
  ! does this (the reference's quirk, kept)
  real(r8) function sumf_areas(sf) result(s)
    type(scalar_field), intent(in) :: sf
    real(r8), pointer :: areag(:), trarea(:), leng(:), dleng(:)
    real(r8) :: sa
    integer :: i
    call mesh_arrays(areag, trarea, leng, dleng)
    s = 0._r8
    sa = 0._r8
    select case (sf%pos)
    case (0)
      do i = 1, sf%n
        sa = sa + areag(i)
        s = s + areag(i)*sf%f(i)
      end do
    case (1)
      do i = 1, sf%n
        sa = sa + trarea(i)
        s = s + trarea(i)*sf%f(i)
      end do
    case (3, 6)
      do i = 1, sf%n
        sa = sa + leng(i)*dleng(i)*0.5_r8
        s = s + areag(i)*sf%f(i)
      end do
    case default
      print*, "Don't know how to calculate sumf_areas on this position:", sf%pos
      stop 1
    end select
    s = s/sa
  end function sumf_areas

  ! sumfsq_areas (src/swm.f90:3442-3490)
  real(r8) function sumfsq_areas(sf) result(s)
    type(scalar_field), intent(in) :: sf
    real(r8), pointer :: areag(:), trarea(:), leng(:), dleng(:)
    real(r8) :: sa, a
    integer :: i
    call mesh_arrays(areag, trarea, leng, dleng)
    s = 0._r8
    sa = 0._r8
    select case (sf%pos)
    case (0)
      do i = 1, sf%n
        sa = sa + areag(i)
        s = s + areag(i)*sf%f(i)**2
      end do
    case (1)
      do i = 1, sf%n
        sa = sa + trarea(i)
        s = s + trarea(i)*sf%f(i)**2
      end do
    case (3, 6)
      do i = 1, sf%n
        a = leng(i)*dleng(i)*0.5_r8
        sa = sa + a
        s = s + a*sf%f(i)**2
      end do
    case default
      print*, "Don't know how to calculate sumfsq_areas on this position:", sf%pos
      stop 1
    end select
    s = s/sa
  end function sumfsq_areas

  ! error_calc (src/swm.f90:3492-3543)
  subroutine error_calc(f, fref, fdif, errormaxrel, error2rel, errormax)
    type(scalar_field), intent(in) :: f, fref
    type(scalar_field), intent(inout) :: fdif
    real(r8) :: errormaxrel, error2rel, errormax
    real(r8) :: refsum, refmax
    fdif = f
    fdif%f = f%f - fref%f
    refsum = sumfsq_areas(fref)
    if (refsum > 0) then
      error2rel = dsqrt(sumfsq_areas(fdif))/dsqrt(refsum)
    else
      error2rel = dsqrt(sumfsq_areas(fdif))
    end if
    refmax = maxval(abs(fref%f(1:fref%n)))
    errormax = maxval(abs(fdif%f(1:fdif%n)))
    if (refmax > 0) then
      errormaxrel = errormax/refmax
    else
      errormaxrel = errormax
    end if
  end subroutine error_calc

  ! calc_energies (src/swm.f90:2686-2752) of the module state h, u, bt, with
  ! the reference's loops on the host; h_ed (trsk) and ke_hx (pxt16) come
  ! from a reference-order tendency of that state
  subroutine calc_energies(Penergy, Kenergy, Tenergy, Availenergy)
    real(r8) :: Penergy, Kenergy, Tenergy, Availenergy
    real(r8), pointer :: areag(:), trarea(:), leng(:), dleng(:)
    real(r8), allocatable :: hed(:), ke(:)
    type(scalar_field) :: hbt
    real(r8) :: mean_height, mhb, unavail
    integer :: i, l
    call mesh_arrays(areag, trarea, leng, dleng)
    allocate(hed(mdesc%ne), ke(mdesc%nv))
    call tendency_fields(h%f, u%f)
    call imx_check(imx_swm_get_field(ref_handle(), IMX_FIELD_H_ED, hed), 'calc_energies')
    call imx_check(imx_swm_get_field(ref_handle(), IMX_FIELD_KE_HX, ke), 'calc_energies')
    Penergy = 0._r8
    do i = 1, mdesc%nv
      Penergy = Penergy + areag(i)*grav*h%f(i)*(h%f(i)*0.5 + bt%f(i))
    end do
    hbt = h
    hbt%f = h%f + bt%f
    mean_height = sumf_areas(hbt)
    unavail = 0._r8
    do i = 1, mdesc%nv
      mhb = mean_height - bt%f(i)
      unavail = unavail + areag(i)*grav*mhb*(mhb*0.5 + bt%f(i))
    end do
    Kenergy = 0._r8
    if (rec_code() == IMX_REC_PERHX) then
      do i = 1, mdesc%nv
        Kenergy = Kenergy + ke(i)*areag(i)
      end do
    else
      do l = 1, mdesc%ne
        Kenergy = Kenergy + 0.5*leng(l)*dleng(l)*hed(l)*u%f(l)**2
      end do
    end if
    Tenergy = Penergy + Kenergy
    Availenergy = Tenergy - unavail
  end subroutine calc_energies

  subroutine mesh_arrays(areag, trarea, leng, dleng)
    real(r8), pointer :: areag(:), trarea(:), leng(:), dleng(:)
    call c_f_pointer(mdesc%hx_areag, areag, [mdesc%nv])
    call c_f_pointer(mdesc%tr_areag, trarea, [mdesc%nt])
    call c_f_pointer(mdesc%ed_leng, leng, [mdesc%ne])
    call c_f_pointer(mdesc%edhx_leng, dleng, [mdesc%ne])
  end subroutine mesh_arrays

  ! The handle whose fields are the reference's: the run's own with
  ! IMX_PREC_FP64_REF, else the FP64 reference-order companion
  type(c_ptr) function ref_handle() result(p)
    if (precision_code() == IMX_PREC_FP64_REF) then
      p = handle
    else
      if (.not. c_associated(diag_handle)) then
        call open_handle(diag_handle, IMX_PREC_FP64_REF)
        call imx_check(imx_swm_set_topography(diag_handle, bt%f), 'ref_handle')
      end if
      p = diag_handle
    end if
  end function ref_handle

  ! a reference-order tendency of (hh, uu), leaving its intermediate fields
  ! on ref_handle()
  subroutine tendency_fields(hh, uu)
    real(r8), intent(in) :: hh(:), uu(:)
    real(r8), allocatable :: m(:), f(:)
    allocate(m(mdesc%nv), f(mdesc%ne))
    call imx_check(imx_swm_tendency(ref_handle(), hh, uu, m, f), 'tendency_fields')
  end subroutine tendency_fields

  ! The fields the time loop's diagnostics read: with IMX_PREC_FP64_REF
  ! those its last RK stage left (the reference's reading), else a
  ! reference-order tendency of the handle's state
  subroutine ref_fields(fresh)
    logical, intent(in) :: fresh
    if (fresh) then
      call imx_check(imx_swm_get_state(handle, h%f, u%f), 'ref_fields')
      call tendency_fields(h%f, u%f)
    end if
  end subroutine ref_fields

  ! pxt16: calc_energies takes K as sum ke_hx A (src/swm.f90:2733-2737), the
  ! Perot kinetic energy of the fields the reference reads; the device
  ! diagnostics hold 1/2 l d h_e u^2. Rewrites K, T and A of d.
  subroutine perot_kenergy(d, initial)
    real(c_double), intent(inout) :: d(:)
    logical, intent(in) :: initial
    real(r8), allocatable :: ke(:)
    real(r8), pointer :: areag(:), trarea(:), leng(:), dleng(:)
    real(r8) :: kp, unavail
    integer :: i
    if (rec_code() /= IMX_REC_PERHX) return
    call mesh_arrays(areag, trarea, leng, dleng)
    allocate(ke(mdesc%nv))
    ! the reference's first energies follow a tendency at the initial state
    call ref_fields(initial .or. precision_code() /= IMX_PREC_FP64_REF)
    call imx_check(imx_swm_get_field(ref_handle(), IMX_FIELD_KE_HX, ke), 'perot_kenergy')
    kp = 0._r8
    do i = 1, mdesc%nv
      kp = kp + ke(i)*areag(i)
    end do
    unavail = d(4) - d(5)
    d(3) = kp
    d(4) = d(2) + kp
    d(5) = d(4) - unavail
  end subroutine perot_kenergy

  subroutine close_handles()
    if (c_associated(handle)) call imx_swm_destroy(handle)
    if (c_associated(diag_handle)) call imx_swm_destroy(diag_handle)
    handle = c_null_ptr
    diag_handle = c_null_ptr
  end subroutine close_handles

  ! swmname and meshname as swmpars builds them (src/swm_data.f90:640-683);
  ! without the time step when swmpars runs with usetime = .false. (the LTE
  ! analysis)
  subroutine make_names(glevel, usetime)
    integer, intent(in) :: glevel
    logical, intent(in) :: usetime
    character(len=16) :: atmp
    character(len=32) :: dtpart
    write(atmp, '(i8)') testcase
    swmname = "swm_tc"//trim(adjustl(atmp))
    write(atmp, '(i8)') nint(dt)
    dtpart = ""
    if (usetime) dtpart = "_dt"//trim(adjustl(atmp))
    ! stag, wrapper (or the explicit method names), area method (swm_data.f90:640-656)
    if (trim(mtdwrapper) /= 'none') then
      swmname = trim(swmname)//trim(dtpart)//"_"//trim(stag)//"_"//trim(adjustl(mtdwrapper))// &
                "_areageo"
    else
      swmname = trim(swmname)//trim(dtpart)//"_"//trim(stag)//"_vrec"//trim(adjustl(reconmtd))
      if (rec_code() == IMX_REC_GASS) then
        write(atmp, '(f4.2)') real(gass_coef())
        swmname = trim(swmname)//trim(adjustl(trim(atmp)))
      end if
      swmname = trim(swmname)//"_crec"//trim(adjustl(coriolis_reconmtd))//"_sint"// &
                trim(adjustl(sinterpol))//"_grd"//trim(adjustl(gradmtd))//"_areageo"
    end if
    ! scheme-option suffixes of swmpars (src/swm_data.f90:658-683)
    if (trim(pv_stab) /= 'none') then
      write(atmp, '(f5.2)') real(pvspar)
      swmname = trim(swmname)//"_pvs"//trim(adjustl(trim(pv_stab)))//trim(adjustl(trim(atmp)))
    end if
    if (inopv > 0) swmname = trim(swmname)//"_nopv"
    if (hollgw > 0 .and. testcase >= 32 .and. testcase <= 35) then
      write(atmp, '(f6.2)') real(hollgw)
      swmname = trim(swmname)//"_hol"//trim(adjustl(trim(atmp)))
    end if
    if (k2_max > 0) then
      write(atmp, '(f10.3)') real(dlog(k2_max)/dlog(10._r8))
      swmname = trim(swmname)//"_"//trim(diffus)//"_diffusion_10to"//trim(adjustl(trim(atmp)))
    end if
    if (k4_max > 0) then
      write(atmp, '(f10.3)') real(dlog(k4_max)/dlog(10._r8))
      swmname = trim(swmname)//"_"//trim(hyperdiffus)//"_hyperdiffusion_10to"//trim(adjustl(trim(atmp)))
    end if
    ! the grid's name (namegrid) ends every output file name
    meshname = mesh%name

  end subroutine make_names

  ! One libimx handle with the options of par/swm.par (precision prec)
  subroutine open_handle(hnd, prec)
    type(c_ptr), intent(inout) :: hnd
    integer(c_int32_t), intent(in) :: prec
    type(imx_swm_opts) :: o
    type(imx_scheme) :: sc
    integer(c_int) :: rc
    call imx_swm_default_opts(o)
    o%precision = prec
    o%testcase = merge(2, testcase, testcase == 100)
    o%use_sfc = use_sfc
    o%use_graph = use_graph
    o%erad = erad
    o%grav = grav
    o%omega = omega
    o%dt = dt
    call imx_check(imx_swm_create(mdesc, o, hnd), 'imx_swm_create')
    ! the f-sphere of the test case, as initialize_fields sets it
    block
      integer(c_int32_t) :: fs
      real(c_double) :: fc
      call imx_check(imx_ic_testcase_coriolis(int(testcase, c_int32_t), fs, fc), 'open_handle')
      fsphere = fs
      fcte = fc
      if (fs /= 0) call imx_check(imx_swm_set_coriolis(hnd, fs, fc), 'imx_swm_set_coriolis')
    end block
    if (scheme_options()) then
      sc%stag = merge(1, 0, trim(stag) == 'HC')
      sc%wrap = wrap_code()
      sc%apvm = merge(1, 0, trim(pv_stab) == 'apvm')
      sc%pvspar = pvspar
      sc%k2 = max(k2_max, 0._r8)
      sc%diffus = max(coef_code(diffus), 0)
      sc%k4 = max(k4_max, 0._r8)
      sc%hyperdiffus = max(coef_code(hyperdiffus), 0)
      sc%sinterp = sint_code()
      sc%coriolis = cor_code()
      sc%recon = rec_code()
      sc%gasscoef = gass_coef()
      sc%nopv = merge(1, 0, inopv > 0)
      ! the fused fp32/fp64 kernels carry the wrappers, most method
      ! triples and the level model with trsk10; the library names what
      ! they refuse
      rc = imx_swm_set_scheme(hnd, mdesc, sc)
      if (rc == IMX_ENOTSUP .and. prec /= IMX_PREC_FP64_REF) then
        write(*, '(a)') " ERROR: these scheme options run on the fp64_ref path (par/swm_gpu.par):"
        write(*, '(a)') "        "//trim(imx_error_message())
        stop 1
      end if
      call imx_check(rc, 'imx_swm_set_scheme')
    end if
  end subroutine open_handle

  ! swm_horiz_loc_trunc_er (src/swm.f90:382-734): TC2's closed forms against
  ! one tendency evaluation (imx_swm_lte: the tendency and its analysis fields
  ! on the GPU). Prints the reference's table and mimetic sums and appends
  ! the table to data/swm_loc_trunc_errors.txt in its layout.
  subroutine swm_horiz_loc_trunc_er(meshtmp)
    type(grid_structure), intent(in) :: meshtmp
    real(c_double) :: err(3*IMX_LTE_NROW), mim(IMX_LTE_NMIM)
    character(len=16), parameter :: rows(IMX_LTE_NROW) = [character(len=16) :: &
      " h_ed ", " divuh ", " kenergy ", " abs_vort ", " h_tr ", " q_tr ", " q_hx ", " q_ed ", &
      " grad_ghbK ", " grad_ghbK_exact ", " vhq_tr ", " q_grad_tr ", " vh_hx ", " vhq_hx ", &
      " uhq_perp ", " lapu ", " masseq ", " momeq "]
    character(len=100), parameter :: fmt = "(a16, i12,  3e18.8, a80)"
    integer :: fu, r
    logical :: ex

    mesh = meshtmp
    call swmpars(usetime=.false.)
    if (testcase /= 2 .and. (testcase < 32 .or. testcase > 35)) then
      write(*, '(a)') " ERROR: the local truncation error analysis has the closed forms of test cases 2 and 32-35"
      stop 1
    end if
    if (trim(precision) /= 'fp64_ref') then
      write(*, '(a)') " ERROR: the local truncation error analysis runs on the fp64_ref path (par/swm_gpu.par)"
      stop 1
    end if
    call allocate_globalswmvars()
    call initialize_gridprop()
    call make_names(mesh%glevel, .false.)
    call imx_check(imx_swm_lte_hol(handle, mesh%lib, hollgw, err, mim), 'swm_horiz_loc_trunc_er')
    call close_handles()

    call imx_check(imx_make_dir('data'//c_null_char), 'swm_horiz_loc_trunc_er')
    inquire(file="data/swm_loc_trunc_errors.txt", exist=ex)
    if (ex) then
      open(newunit=fu, file="data/swm_loc_trunc_errors.txt", status='old', position='append')
    else
      open(newunit=fu, file="data/swm_loc_trunc_errors.txt", status='replace')
      write(fu, *) "          Field       Mesh       MaxErrorRel       "//&
        " RMSError         MaxError                            "//&
        "Methods                                     Grid"
    end if
    print*, " Field       Mesh       MaxErrorRel        RMSError         MaxError"
    do r = 1, IMX_LTE_NROW
      print '(a16, i12,  3e18.8)', rows(r), mdesc%nv, err(3*r-2), err(3*r-1), err(3*r)
      write(fu, trim(fmt)) rows(r), mdesc%nv, err(3*r-2), err(3*r-1), err(3*r), &
        trim(swmname)//" "//trim(meshname)
    end do
    close(fu)
    print*
    print*, "Verify mimetic properties (normalized)"
    print*, "-------------------------"
    print*
    print*, "1) Mass Conservation"
    print*, " To be analysed on the evolving system"
    print*
    print*, "2) Product rule - normalized fields"
    print*, "Sum of h*Div(u): ", mim(1)
    print*, "Sum of h*Div(u)+u*Grad(h): ", mim(2)
    print*, "Sum of Div(hu): ", mim(3)
    print*, "    Difference:", mim(4)
    print*
    print*, "3) Cross rule - normalized fields"
    print*, "Sum of k.Rot(Grad(h)):", mim(5)
    print*
    print*, "4) Coriolis spurious energy budget - normalized fields "
    print*, "Sum of Coriolis energy (Ae*q*h^2*u*u_perp):", mim(6)
    print*, "Sum of Coriolis energy / Tenergy:", mim(7)
  end subroutine swm_horiz_loc_trunc_er

  ! swm_normal_mode_analysis (src/swm.f90:738-976): the linearised operator
  ! of the initial state by one-sided differences with pert = mean cell
  ! area, written as the reference does: dense rows for n < 100, else the
  ! (i, j, M(i,j)) triplets with |M| > epsilon(1.)/100. The library
  ! evaluates one tendency per colour of columns with disjoint stencils;
  ! every entry equals the one-column difference bit for bit.
  subroutine swm_normal_mode_analysis()
    real(r8), pointer :: areag(:)
    integer(c_int32_t), allocatable :: row(:), col(:)
    real(r8), allocatable :: val(:), dense(:, :)
    integer(c_int64_t) :: nnz
    integer(c_int32_t) :: ncolors
    real(r8) :: pert
    integer :: n, k, iunit
    character(len=512) :: filename
    character(len=100) :: fmt, atmp
    call c_f_pointer(mdesc%hx_areag, areag, [mdesc%nv])
    pert = sum(areag)/mdesc%nv            ! mesh%meanhxarea
    n = mdesc%nv + mdesc%ne
    print*, "Generating the linearised operator (", n, " unknowns)"
    call imx_check(imx_swm_jacobian(handle, mdesc, h%f, u%f, pert, 0_c_int64_t, [0_c_int32_t], [0_c_int32_t], &
                                    [0._r8], nnz, ncolors), 'swm_normal_mode_analysis')
    allocate(row(nnz), col(nnz), val(nnz))
    call imx_check(imx_swm_jacobian(handle, mdesc, h%f, u%f, pert, nnz, row, col, val, nnz, ncolors), &
                   'swm_normal_mode_analysis')
    print*, "Tendency evaluations:", ncolors + 1, " (one per unknown in the reference:", n, ")"
    call imx_check(imx_make_dir('data'//c_null_char), 'swm_normal_mode_analysis')
    if (n < 100) then
      filename = "data/"//trim(swmname)//"_linearmatrix_"//trim(meshname)//".txt"
      allocate(dense(n, n))
      dense = 0._r8
      do k = 1, int(nnz)
        dense(row(k), col(k)) = val(k)
      end do
      write(atmp, '(i10)') n
      fmt = '('//trim(adjustl(atmp))//'f24.8)'
      open(newunit=iunit, file=trim(filename), status='replace')
      do k = 1, n
        write(iunit, fmt) dense(k, 1:n)
      end do
    else
      filename = "data/"//trim(swmname)//"_linearmatrixsparse_"//trim(meshname)//".txt"
      open(newunit=iunit, file=trim(filename), status='replace')
      do k = 1, int(nnz)
        if (abs(val(k)) > epsilon(1.)/100._r8) write(iunit, '(2i10, es26.16e3)') row(k), col(k), val(k)
      end do
    end if
    close(iunit)
    print*, "Matrix printed as file: ", trim(filename)
  end subroutine swm_normal_mode_analysis

  ! plotfields (src/swm.f90:2754-2904) for runs without the local
  ! truncation-error test: at k = 0, every plotsteps steps and at the end,
  ! h, eta, K, q_tr (pv), q_ed (pv_ed), div(uh) and, with the analytic
  ! solution, the h, u and q_tr errors, each sampled on the 1440 x 720
  ! lat-lon grid by imx_remap_ll (plot_scalarfield). eta, K, q_ed and div(uh)
  ! are intermediates of the last RK stage that only IMX_PREC_FP64_REF keeps;
  ! fused handles write h, pv and the errors.
  subroutine plotfields(k, time)
    integer, intent(in) :: k
    real(r8), intent(in) :: time
    character(len=60) :: atime
    real(r8), allocatable :: f(:)
    logical :: ref
    if (nplots <= 0) return
    if (.not. (k == ntime .or. mod(k, plotsteps) == 0)) return
    write(atime, *) nint(time)
    atime = adjustl(atime)
    ref = precision_code() == IMX_PREC_FP64_REF
    call imx_check(imx_swm_get_state(handle, h%f, u%f), 'plotfields')
    call plot_scalarfield(h%f, 0, "_h_t"//trim(atime))
    if (ref) then
      allocate(f(max(mdesc%nv, mdesc%ne, mdesc%nt)))
      call imx_check(imx_swm_get_field(handle, IMX_FIELD_ETA, f), 'plotfields')
      call plot_scalarfield(f(1:mdesc%nt), 1, "_eta_t"//trim(atime))
      call imx_check(imx_swm_get_field(handle, IMX_FIELD_KE_HX, f), 'plotfields')
      call plot_scalarfield(f(1:mdesc%nv), 0, "_Kenergy_t"//trim(atime))
    end if
    if (.not. allocated(qtr)) allocate(qtr(mdesc%nt))
    call imx_check(imx_swm_get_field(handle, IMX_FIELD_Q_TR, qtr), 'plotfields')
    call plot_scalarfield(qtr, 1, "_pv_t"//trim(atime))
    if (ref) then
      call imx_check(imx_swm_get_field(handle, IMX_FIELD_Q_ED, f), 'plotfields')
      call plot_scalarfield(f(1:mdesc%ne), 6, "_pv_ed_t"//trim(atime))
      call imx_check(imx_swm_get_field(handle, IMX_FIELD_DIVUH, f), 'plotfields')
      call plot_scalarfield(f(1:mdesc%nv), 0, "_divuh_t"//trim(atime))
    end if
    if (refsol) then
      call plot_scalarfield(h%f - h_exact%f, 0, "_h_error_t"//trim(atime))
      call plot_scalarfield(u%f - u_exact%f, 6, "_u_error_t"//trim(atime))
      call plot_scalarfield(qtr - q_exact, 1, "_q_tr_error_t"//trim(atime))
    end if
  end subroutine plotfields

  ! plot_scalarfield (src/interpack.f90:7376-7514): 720 x 1440 float32
  ! (lon, lat, value) stream in data/<swmname><suffix>_<grid>.dat
  subroutine plot_scalarfield(f, pos, suffix)
    real(r8), intent(in) :: f(:)
    integer, intent(in) :: pos
    character(len=*), intent(in) :: suffix
    integer, parameter :: nlat = 720, nlon = 1440
    real(c_float), allocatable :: buf(:)
    real(r8), allocatable :: fc(:)
    character(len=512) :: filename
    integer :: iunit
    allocate(buf(3*nlat*nlon), fc(size(f)))
    fc = f
    call imx_check(imx_remap_ll(mdesc, int(pos, c_int32_t), fc, nlon, nlat, buf), 'plot_scalarfield')
    call imx_check(imx_make_dir('data'//c_null_char), 'plot_scalarfield')
    filename = "data/"//trim(swmname)//trim(suffix)//"_"//trim(meshname)//".dat"
    open(newunit=iunit, file=trim(filename), status='replace', access='stream', form='unformatted')
    write(iunit) buf
    close(iunit)
  end subroutine plot_scalarfield

  ! error_calc (src/swm.f90:3492-3543) with sumfsq_areas weights w
  ! (:3442-3490) on plain arrays: relative max and L2 errors of f against fref
  subroutine error_calc_w(f, fref, w, errormaxrel, error2rel, errormax)
    real(r8), intent(in) :: f(:), fref(:), w(:)
    real(r8), intent(out) :: errormaxrel, error2rel, errormax
    real(r8) :: refsum, refmax, difsum, sw
    integer :: i
    refsum = 0._r8; difsum = 0._r8; sw = 0._r8
    do i = 1, size(f)
      sw = sw + w(i)
      refsum = refsum + w(i)*fref(i)**2
      difsum = difsum + w(i)*(f(i) - fref(i))**2
    end do
    refsum = refsum/sw
    difsum = difsum/sw
    if (refsum > 0) then
      error2rel = sqrt(difsum)/sqrt(refsum)
    else
      error2rel = sqrt(difsum)
    end if
    refmax = maxval(abs(fref))
    errormax = maxval(abs(f - fref))
    if (refmax > 0) then
      errormaxrel = errormax/refmax
    else
      errormaxrel = errormax
    end if
  end subroutine error_calc_w

  ! The analytic-solution branch of the time loop (src/swm.f90:291-349):
  ! errors of h, u and q_tr against the exact fields, max |div u| and
  ! max |grad K| over edges, written by write_evol_error every nprints
  ! steps and by write_errors at the last step. div u and K are the
  ! reference's div_hx and kinetic_energy_hx for the run's scheme (HC signs,
  ! Perot K under pxt16): with IMX_PREC_FP64_REF those of the last RK stage,
  ! as the reference reads them; fused handles evaluate them on the state
  ! with an FP64 reference-order tendency (ref_fields).
  ! Returns .false. when the run should stop (blow-up, swm.f90:331-340).
  logical function write_ref_errors(k, time, d, d0) result(go_on)
    integer, intent(in) :: k
    real(r8), intent(in) :: time, d(:), d0(:)
    real(r8), pointer :: areag(:), leng(:), dleng(:), trarea(:)
    real(r8), allocatable :: we(:), divu(:), ke(:)
    real(r8) :: emr_h, e2_h, em_h, emr_u, e2_u, em_u, emr_q, e2_q, em_q
    real(r8) :: maxdiv, max_gradke, rmsdiv, sw
    integer(c_int32_t), pointer :: ev(:)
    integer :: l
    go_on = .true.
    call c_f_pointer(mdesc%hx_areag, areag, [mdesc%nv])
    call c_f_pointer(mdesc%ed_leng, leng, [mdesc%ne])
    call c_f_pointer(mdesc%edhx_leng, dleng, [mdesc%ne])
    call c_f_pointer(mdesc%tr_areag, trarea, [mdesc%nt])
    call c_f_pointer(mdesc%ed_v, ev, [2*mdesc%ne])
    call imx_check(imx_swm_get_state(handle, h%f, u%f), 'imx_swm_get_state')
    call imx_check(imx_swm_get_field(handle, IMX_FIELD_Q_TR, qtr), 'imx_swm_get_field')
    allocate(we(mdesc%ne), divu(mdesc%nv), ke(mdesc%nv))
    we = leng*dleng*0.5_r8
    call error_calc_w(h%f, h_exact%f, areag, emr_h, e2_h, em_h)
    call error_calc_w(u%f, u_exact%f, we, emr_u, e2_u, em_u)
    call error_calc_w(qtr, q_exact, trarea, emr_q, e2_q, em_q)
    call ref_fields(precision_code() /= IMX_PREC_FP64_REF)
    call imx_check(imx_swm_get_field(ref_handle(), IMX_FIELD_DIVU, divu), 'imx_swm_get_field')
    call imx_check(imx_swm_get_field(ref_handle(), IMX_FIELD_KE_HX, ke), 'imx_swm_get_field')
    sw = sum(areag)
    rmsdiv = sqrt(sum(areag*divu**2)/sw)
    maxdiv = maxval(abs(divu))
    max_gradke = 0._r8
    do l = 1, mdesc%ne   ! HTC: ed%v(1) -> ed%v(2) (swm.f90:318-326)
      max_gradke = max(max_gradke, abs((ke(ev(2*l)) - ke(ev(2*l-1)))/leng(l)/erad))
    end do
    if (mod(k, nprints) == 0) &
      call write_evol_error(k, time, (d(1) - d0(1))/d0(1), emr_h, e2_h, emr_u, e2_u, emr_q, e2_q, &
                            (d(2) - d0(2))/d0(2), (d(3) - d0(3))/d0(3), (d(4) - d0(4))/d0(4), &
                            (d(5) - d0(5))/d0(5), rmsdiv, maxdiv, max_gradke, 1._r8)
    if ((emr_h > 20.0_r8 .or. isnan(emr_h)) .and. k > 3) then
      blowup = blowup + 1
      print*, "System might be unstable, large errors:", emr_h
      if (blowup >= 5 .or. isnan(emr_h) .or. emr_h > 100000000.0_r8) then
        print*, "Stopping due to large errors", emr_h
        go_on = .false.
        return
      end if
    end if
    if (k == ntime) &
      call write_errors(time, emr_h, e2_h, emr_u, e2_u, (d(1) - d0(1))/d0(1), (d(2) - d0(2))/d0(2), &
                        (d(3) - d0(3))/d0(3), (d(4) - d0(4))/d0(4), 1._r8)
  end function write_ref_errors

  ! write_evol_error (src/swm.f90:2999-3086): same file name, header, format
  subroutine write_evol_error(k, time, tmass, errormax_h, error2_h, errormax_u, error2_u, &
                              errormax_pv, error2_pv, penergy, kenergy, tenergy, availenergy, &
                              rmsdiv, maxdiv, max_gradke, tmp)
    integer, intent(in) :: k
    real(r8), intent(in) :: time, tmass, errormax_h, error2_h, errormax_u, error2_u, errormax_pv, &
                            error2_pv, penergy, kenergy, tenergy, availenergy, rmsdiv, maxdiv, &
                            max_gradke, tmp
    character(len=512) :: filename, buffer
    integer, save :: iunit = 0
    logical, save :: opened = .false.
    filename = "data/"//trim(swmname)//"_evolution_"//trim(meshname)//".txt"
    if (.not. opened) then
      opened = .true.
      call imx_check(imx_make_dir('data'//c_null_char), 'write_evol_error')
      buffer = "        n        mvdist  tcase       k       nt    "// &
               "     dt                 cfl               TimeDays  "// &
               "        errormax_h       error2_h               errormax_u     "// &
               "   error2_u               errormax_pv       error2_pv            "// &
               " mass              Penergy      "// &
               "     Kenergy             Tenergy          Availenergy      "// &
               "     RMSdiv              maxdiv    "// &
               "     max_gradke          tmp        SWMparameters"
      open(newunit=iunit, file=trim(filename), status='replace')
      write(iunit, '(a)') trim(buffer)
    end if
    write(iunit, '(i12, 1f12.4, i4, 2i10, 18e20.10 , a100)') mdesc%nv, meanvdist*rad2deg, testcase, &
      k, ntime, dt, cfl, time*sec2day, errormax_h, error2_h, errormax_u, error2_u, errormax_pv, &
      error2_pv, tmass, penergy, kenergy, tenergy, availenergy, rmsdiv, maxdiv, max_gradke, tmp, &
      trim(adjustl(trim(swmname)))//"_"//trim(adjustl(trim(meshname)))
    flush(iunit)
  end subroutine write_evol_error

  ! write_errors (src/swm.f90:2922-2997): appended to data/swm_errors.txt
  subroutine write_errors(time, errormax_h, error2_h, errormax_u, error2_u, error_mass, &
                          penergy, kenergy, tenergy, tmp)
    real(r8), intent(in) :: time, errormax_h, error2_h, errormax_u, error2_u, error_mass, &
                            penergy, kenergy, tenergy, tmp
    character(len=512) :: buffer
    character(len=*), parameter :: filename = "data/swm_errors.txt"
    integer :: iunit
    logical :: ifile
    buffer = "       Mesh  TCase    nTime"// &
             "     dt                   TimeDays  "// &
             "        MaxError_h          RMSError_h            MaxError_u     "// &
             "   RMSError_u               Mass              Penergy      "// &
             "     Kenergy             Tenergy             ExtraData                    "// &
             "                Methods  Grid    "
    call imx_check(imx_make_dir('data'//c_null_char), 'write_errors')
    inquire(file=filename, exist=ifile)
    if (ifile) then
      open(newunit=iunit, file=filename, status='old', position='append')
    else
      open(newunit=iunit, file=filename, status='replace')
      write(iunit, '(a)') buffer
    end if
    write(iunit, '(i12, i4, i10, 11e20.10 , a120)') mdesc%nv, testcase, ntime, dt, time*sec2day, &
      errormax_h, error2_h, errormax_u, error2_u, error_mass, penergy, kenergy, tenergy, tmp, &
      trim(adjustl(trim(swmname)))//" "//trim(adjustl(trim(meshname)))
    close(iunit)
  end subroutine write_errors

  ! write_evol (src/swm.f90:3088-3152): same file name, header and format
  subroutine write_evol(k, time, tmass, penergy, kenergy, tenergy, availenergy, rmsdiv)
    integer, intent(in) :: k
    real(r8), intent(in) :: time, tmass, penergy, kenergy, tenergy, availenergy, rmsdiv
    character(len=512) :: filename, buffer
    integer, save :: iunit = 0
    logical, save :: opened = .false.
    filename = "data/"//trim(swmname)//"_mass_energy_"//trim(meshname)//".txt"
    if (.not. opened) then
      opened = .true.
      call imx_check(imx_make_dir('data'//c_null_char), 'write_evol')
      buffer = "        n        mvdist  tcase       k       nt        dt(s)       "// &
               "          cfl                 time(dys)             mass               Penergy    "// &
               "          Kenergy            Tenergy            Availenergy           RMSdiv"
      open(newunit=iunit, file=trim(filename), status='replace')
      write(iunit, '(a)') trim(buffer)
    end if
    write(iunit, '(i12, 1f12.4, i4, 2i10, 9e28.16)') mdesc%nv, meanvdist*rad2deg, testcase, &
      k, ntime, dt, cfl, time/day2sec, tmass, penergy, kenergy, tenergy, availenergy, rmsdiv
    flush(iunit)
    write(*, '(i12, 1f8.3, i4, 2i10, 2f10.4, 6e16.6)') mdesc%nv, meanvdist*rad2deg, testcase, &
      k, ntime, dt, time/day2sec, tmass, penergy, kenergy, tenergy, availenergy, rmsdiv
  end subroutine write_evol

  ! final (h, u) in caller numbering as IMXBIN records "h" and "u"
  subroutine dump_final_state()
    integer :: fu
    character(len=24) :: nm
    open(newunit=fu, file="data/"//trim(swmname)//"_state_"//trim(meshname)//".bin", &
         access='stream', form='unformatted', status='replace')
    write(fu) 'IMXBIN01'
    nm = 'h'
    write(fu) nm, 'f8  ', int(size(h%f), 8), h%f
    nm = 'u'
    write(fu) nm, 'f8  ', int(size(u%f), 8), u%f
    close(fu)
  end subroutine dump_final_state
end module swm
