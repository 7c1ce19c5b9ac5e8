!=======================================================================
! swm: the shallow-water module of the imodel driver, on the GPU.
! Keeps the reference call surface (src/swm.f90): swm_tests (:102-380),
! swm_horiz_loc_trunc_er (:382-734),
! tendency(h, u, masseq, momeq) (:2416-2598), ode_rk4 (:2600-2683); the
! state lives in HBM behind one libimx handle (the reference keeps it in
! the swm_data module globals, src/swm_data.f90:31-334).
!=======================================================================
module swm
  use iso_c_binding
  use imx_c
  use swm_par
  implicit none
  real(r8), parameter :: erad = 6.37122e6_r8   ! constants.f90:55
  real(r8), parameter :: grav = 9.80616_r8     ! constants.f90:61
  real(r8), parameter :: omega = 7.292e-5_r8   ! constants.f90:67
  real(r8), parameter :: rad2deg = 57.295779513082323_r8

  real(r8), parameter :: sec2day = 1._r8/86400._r8
  type(c_ptr) :: handle = c_null_ptr
  type(imx_mesh_desc) :: mdesc
  real(r8), allocatable :: h(:), u(:)
  ! analytic reference solution (RefSolAnal, src/swm_data.f90:685): TC2
  logical :: refsol = .false.
  integer :: blowup = 0, plotsteps = 1
  real(r8), allocatable :: h_exact(:), u_exact(:), q_exact(:), qtr(:)
  character(len=256) :: swmname = '', meshname = ''
  real(r8) :: cfl = 0._r8, meanvdist = 0._r8

contains

  subroutine swm_tests(meshp, glevel)
    type(c_ptr), intent(in) :: meshp
    integer, intent(in) :: glevel
    real(c_double) :: d0(IMX_NDIAG), d(IMX_NDIAG)
    real(r8), pointer :: leng(:)
    real(r8) :: maxvel, minvdist, time
    integer :: k, kn

    call read_swm_par(glevel)
    call imx_check(imx_mesh_desc_of(meshp, mdesc), 'imx_mesh_desc_of')
    allocate(h(mdesc%nv), u(mdesc%ne))
    if (testcase == 2) then
      if (trim(stag) == 'HC') then
        call imx_check(imx_ic_tc2_hc(meshp, erad, h, u), 'imx_ic_tc2_hc')
      else
        call imx_check(imx_ic_tc2(meshp, erad, h, u), 'imx_ic_tc2')
      end if
    else
      ! seeded spectral state (new test case 100): h' 5% of H0, 20 m/s rms wind
      call imx_check(imx_ic_spectral(mdesc, int(seed, c_int64_t), 16_c_int32_t, 2.94e4_r8/grav, &
                                     0.05_r8, 20._r8, erad, h, u), 'imx_ic_spectral')
    end if
    maxvel = maxval(abs(u))
    call c_f_pointer(mdesc%ed_leng, leng, [mdesc%ne])
    minvdist = minval(leng)
    meanvdist = sum(leng)/mdesc%ne
    cfl = abs(maxvel*dt/(minvdist*erad))   ! swm.f90:2383
    print*, "CFL:", cfl

    call make_names(glevel, .true.)

    call open_handle()
    call imx_check(imx_swm_set_state(handle, h, u), 'imx_swm_set_state')
    if (normal_modes == 1) then
      call swm_normal_mode_analysis()
      call imx_swm_destroy(handle)
      handle = c_null_ptr
      return
    end if
    call imx_check(imx_swm_diagnostics(handle, d0), 'imx_swm_diagnostics')
    refsol = (testcase == 2)
    if (refsol) then
      allocate(h_exact(mdesc%nv), u_exact(mdesc%ne), q_exact(mdesc%nt), qtr(mdesc%nt))
      h_exact = h
      u_exact = u
      call imx_check(imx_ic_tc2_pv(meshp, erad, q_exact), 'imx_ic_tc2_pv')
    end if
    ! swm.f90:226-231: both files start with the absolute initial values
    call write_evol_error(0, 0._r8, d0(1), 0._r8, 0._r8, 0._r8, 0._r8, 0._r8, 0._r8, &
                          d0(2), d0(3), d0(4), d0(5), 0._r8, 0._r8, 0._r8, 0._r8)
    call write_evol(0, 0._r8, d0(1), d0(2), d0(3), d0(4), d0(5), 0._r8)
    plotsteps = 1
    if (nplots > 0) plotsteps = max(1, ntime/nplots)   ! swm_data.f90:589-597
    call plotfields(0, 0._r8)

    ! time loop of swm_tests (swm.f90:238-370). With the analytic solution
    ! (TC2) errors are written every nprints steps and at the end
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
        call write_evol(k, time, (d(1) - d0(1))/d0(1), (d(2) - d0(2))/d0(2), (d(3) - d0(3))/d0(3), &
                        (d(4) - d0(4))/d0(4), (d(5) - d0(5))/d0(5), d(7))
      end if
      call plotfields(k, time)
    end do

    call imx_check(imx_swm_get_state(handle, h, u), 'imx_swm_get_state')
    if (dump_state == 1) call dump_final_state()
    call imx_swm_destroy(handle)
    handle = c_null_ptr
  end subroutine swm_tests

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
      swmname = trim(swmname)//trim(dtpart)//"_"//trim(stag)// &
                "_vrectrsk_crectrsk_sinttrsk_grdtrsk_areageo"
    end if
    ! scheme-option suffixes of swmpars (src/swm_data.f90:658-683)
    if (trim(pv_stab) /= 'none') then
      write(atmp, '(f5.2)') real(pvspar)
      swmname = trim(swmname)//"_pvs"//trim(adjustl(trim(pv_stab)))//trim(adjustl(trim(atmp)))
    end if
    if (k2_max > 0) then
      write(atmp, '(f10.3)') real(dlog(k2_max)/dlog(10._r8))
      swmname = trim(swmname)//"_"//trim(diffus)//"_diffusion_10to"//trim(adjustl(trim(atmp)))
    end if
    if (k4_max > 0) then
      write(atmp, '(f10.3)') real(dlog(k4_max)/dlog(10._r8))
      swmname = trim(swmname)//"_"//trim(hyperdiffus)//"_hyperdiffusion_10to"//trim(adjustl(trim(atmp)))
    end if
    write(atmp, '(i8)') glevel
    meshname = trim(mkind)//"_"//trim(mpos)//"_"//trim(moptm)//"_"//trim(adjustl(atmp))
    if (len_trim(gridname) > 0) meshname = gridname

  end subroutine make_names

  ! One libimx handle for the run, with the scheme options of par/swm.par
  subroutine open_handle()
    type(imx_swm_opts) :: o
    type(imx_scheme) :: sc
    call imx_swm_default_opts(o)
    o%precision = precision_code()
    o%testcase = merge(2, testcase, testcase == 100)
    o%use_sfc = use_sfc
    o%use_graph = use_graph
    o%erad = erad
    o%grav = grav
    o%omega = omega
    o%dt = dt
    call imx_check(imx_swm_create(mdesc, o, handle), 'imx_swm_create')
    if (scheme_options()) then
      ! the fused fp32/fp64 kernels carry trsk10 and (with HTC) pxt16, with
      ! every other option
      if (trim(precision) /= 'fp64_ref' .and. (wrap_code() == 2 .or. &
          (wrap_code() == 1 .and. trim(stag) == 'HC'))) then
        write(*, '(a)') " ERROR: gass18 and HC with pxt16 run on the fp64_ref path"
        write(*, '(a)') "        (par/swm_gpu.par)"
        stop 1
      end if
      sc%stag = merge(1, 0, trim(stag) == 'HC')
      sc%wrap = wrap_code()
      sc%apvm = merge(1, 0, trim(pv_stab) == 'apvm')
      sc%pvspar = pvspar
      sc%k2 = max(k2_max, 0._r8)
      sc%diffus = max(coef_code(diffus), 0)
      sc%k4 = max(k4_max, 0._r8)
      sc%hyperdiffus = max(coef_code(hyperdiffus), 0)
      call imx_check(imx_swm_set_scheme(handle, mdesc, sc), 'imx_swm_set_scheme')
    end if
  end subroutine open_handle

  ! swm_horiz_loc_trunc_er (src/swm.f90:382-734): TC2's closed forms against
  ! one tendency evaluation (imx_swm_lte: the tendency and its analysis fields
  ! on the GPU). Prints the reference's table and mimetic sums and appends
  ! the table to data/swm_loc_trunc_errors.txt in its layout.
  subroutine swm_horiz_loc_trunc_er(meshp, glevel)
    type(c_ptr), intent(in) :: meshp
    integer, intent(in) :: glevel
    real(c_double) :: err(3*IMX_LTE_NROW), mim(IMX_LTE_NMIM)
    character(len=16), parameter :: rows(IMX_LTE_NROW) = [character(len=16) :: &
      " h_ed ", " divuh ", " kenergy ", " abs_vort ", " h_tr ", " q_tr ", " q_hx ", " q_ed ", &
      " grad_ghbK ", " grad_ghbK_exact ", " vhq_tr ", " q_grad_tr ", " vh_hx ", " vhq_hx ", &
      " uhq_perp ", " lapu ", " masseq ", " momeq "]
    character(len=100), parameter :: fmt = "(a16, i12,  3e18.8, a80)"
    integer :: fu, r
    logical :: ex

    call read_swm_par(glevel)
    if (testcase /= 2) then
      write(*, '(a)') " ERROR: the local truncation error analysis has the closed forms of test case 2"
      stop 1
    end if
    if (trim(precision) /= 'fp64_ref') then
      write(*, '(a)') " ERROR: the local truncation error analysis runs on the fp64_ref path (par/swm_gpu.par)"
      stop 1
    end if
    call imx_check(imx_mesh_desc_of(meshp, mdesc), 'imx_mesh_desc_of')
    call make_names(glevel, .false.)
    call open_handle()
    call imx_check(imx_swm_lte(handle, meshp, err, mim), 'swm_horiz_loc_trunc_er')
    call imx_swm_destroy(handle)
    handle = c_null_ptr

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
    call imx_check(imx_swm_jacobian(handle, mdesc, h, u, pert, 0_c_int64_t, [0_c_int32_t], [0_c_int32_t], &
                                    [0._r8], nnz, ncolors), 'swm_normal_mode_analysis')
    allocate(row(nnz), col(nnz), val(nnz))
    call imx_check(imx_swm_jacobian(handle, mdesc, h, u, pert, nnz, row, col, val, nnz, ncolors), &
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

  ! subroutine tendency(h, u, masseq, momeq)  (src/swm.f90:2416)
  subroutine tendency(hh, uu, masseq, momeq)
    real(r8), intent(in) :: hh(:), uu(:)
    real(r8), intent(inout) :: masseq(:), momeq(:)
    call imx_check(imx_swm_tendency(handle, hh, uu, masseq, momeq), 'tendency')
  end subroutine tendency

  ! subroutine ode_rk4(t, h, u, h_new, u_new, dt)  (src/swm.f90:2600)
  subroutine ode_rk4(hh, uu, h_new, u_new)
    real(r8), intent(in) :: hh(:), uu(:)
    real(r8), intent(inout) :: h_new(:), u_new(:)
    call imx_check(imx_swm_set_state(handle, hh, uu), 'ode_rk4')
    call imx_check(imx_swm_step(handle, 1_c_int32_t), 'ode_rk4')
    call imx_check(imx_swm_get_state(handle, h_new, u_new), 'ode_rk4')
  end subroutine ode_rk4

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
    call imx_check(imx_swm_get_state(handle, h, u), 'plotfields')
    call plot_scalarfield(h, 0, "_h_t"//trim(atime))
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
      call plot_scalarfield(h - h_exact, 0, "_h_error_t"//trim(atime))
      call plot_scalarfield(u - u_exact, 6, "_u_error_t"//trim(atime))
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
  ! (:3442-3490): relative max and L2 errors of f against fref
  subroutine error_calc(f, fref, w, errormaxrel, error2rel, errormax)
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
  end subroutine error_calc

  ! The analytic-solution branch of the time loop (src/swm.f90:291-349):
  ! errors of h, u and q_tr against the exact fields, max |div u| and
  ! max |grad K| over edges, written by write_evol_error every nprints
  ! steps and by write_errors at the last step. With IMX_PREC_FP64_REF,
  ! q_tr, div u and K are those the reference reads (its last RK stage);
  ! fused handles give q_tr of the last stage and div u, K of the state.
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
    call imx_check(imx_swm_get_state(handle, h, u), 'imx_swm_get_state')
    call imx_check(imx_swm_get_field(handle, IMX_FIELD_Q_TR, qtr), 'imx_swm_get_field')
    allocate(we(mdesc%ne), divu(mdesc%nv), ke(mdesc%nv))
    we = leng*dleng*0.5_r8
    call error_calc(h, h_exact, areag, emr_h, e2_h, em_h)
    call error_calc(u, u_exact, we, emr_u, e2_u, em_u)
    call error_calc(qtr, q_exact, trarea, emr_q, e2_q, em_q)
    if (precision_code() == IMX_PREC_FP64_REF) then
      call imx_check(imx_swm_get_field(handle, IMX_FIELD_DIVU, divu), 'imx_swm_get_field')
      call imx_check(imx_swm_get_field(handle, IMX_FIELD_KE_HX, ke), 'imx_swm_get_field')
    else
      call div_ke(u, divu, ke)
    end if
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

  ! div_hx(u) and kinetic_energy_hx (trsk, src/swm_operators.f90:613-661,
  ! :1132-1217) of the given state on the host, caller numbering
  subroutine div_ke(uu, divu, ke)
    real(r8), intent(in) :: uu(:)
    real(r8), intent(out) :: divu(:), ke(:)
    real(r8), pointer :: areag(:), leng(:), dleng(:)
    integer(c_int32_t), pointer :: nnb(:), ved(:), ttg(:)
    real(r8) :: s, k, ca, w
    integer :: i, j, l
    call c_f_pointer(mdesc%hx_areag, areag, [mdesc%nv])
    call c_f_pointer(mdesc%ed_leng, leng, [mdesc%ne])
    call c_f_pointer(mdesc%edhx_leng, dleng, [mdesc%ne])
    call c_f_pointer(mdesc%v_nnb, nnb, [mdesc%nv])
    call c_f_pointer(mdesc%v_ed, ved, [6*mdesc%nv])
    call c_f_pointer(mdesc%hx_ttgout, ttg, [6*mdesc%nv])
    do i = 1, mdesc%nv
      s = 0._r8; k = 0._r8; ca = 0._r8
      do j = 1, nnb(i)
        l = ved(6*(i-1)+j)
        s = s + real(ttg(6*(i-1)+j), r8)*uu(l)*dleng(l)
        w = 0.25_r8*leng(l)*dleng(l)
        ca = ca + w
        k = k + w*uu(l)**2
      end do
      divu(i) = s/areag(i)/erad
      ke(i) = k/ca
    end do
  end subroutine div_ke

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
    write(fu) nm, 'f8  ', int(size(h), 8), h
    nm = 'u'
    write(fu) nm, 'f8  ', int(size(u), 8), u
    close(fu)
  end subroutine dump_final_state
end module swm
