!=======================================================================
! swm: the shallow-water module of the imodel driver, on the GPU.
! Keeps the reference call surface (src/swm.f90): swm_tests (:102-380),
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

  type(c_ptr) :: handle = c_null_ptr
  type(imx_mesh_desc) :: mdesc
  real(r8), allocatable :: h(:), u(:)
  character(len=256) :: swmname = '', meshname = ''
  real(r8) :: cfl = 0._r8, meanvdist = 0._r8

contains

  subroutine swm_tests(meshp, glevel)
    type(c_ptr), intent(in) :: meshp
    integer, intent(in) :: glevel
    type(imx_swm_opts) :: o
    real(c_double) :: d0(IMX_NDIAG), d(IMX_NDIAG)
    real(r8), pointer :: leng(:)
    real(r8) :: maxvel, minvdist, time
    integer :: k, kn
    character(len=16) :: atmp

    call read_swm_par(glevel)
    call imx_check(imx_mesh_desc_of(meshp, mdesc), 'imx_mesh_desc_of')
    allocate(h(mdesc%nv), u(mdesc%ne))
    if (testcase == 2) then
      call imx_check(imx_ic_tc2(meshp, erad, h, u), 'imx_ic_tc2')
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

    write(atmp, '(i8)') testcase
    swmname = "swm_tc"//trim(adjustl(atmp))
    write(atmp, '(i8)') nint(dt)
    swmname = trim(swmname)//"_dt"//trim(adjustl(atmp))//"_HTC_trsk10_areageo"
    write(atmp, '(i8)') glevel
    meshname = trim(mkind)//"_"//trim(mpos)//"_"//trim(moptm)//"_"//trim(adjustl(atmp))

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
    call imx_check(imx_swm_set_state(handle, h, u), 'imx_swm_set_state')
    call imx_check(imx_swm_diagnostics(handle, d0), 'imx_swm_diagnostics')
    call write_evol(0, 0._r8, d0(1), d0(2), d0(3), d0(4), d0(5), d0(7))

    ! time loop of swm_tests (swm.f90:238-370): diagnostics at k<=2,
    ! k==ntime and every nprints steps; the steps in between run on the GPU
    ! without leaving it
    k = 0
    do while (k < ntime)
      if (k < 2) then
        kn = k + 1
      else
        kn = min(ntime, (k/nprints + 1)*nprints)
      end if
      call imx_check(imx_swm_step(handle, int(kn - k, c_int32_t)), 'imx_swm_step')
      k = kn
      time = real(k, r8)*dt
      call imx_check(imx_swm_diagnostics(handle, d), 'imx_swm_diagnostics')
      if (d(8) /= 0._r8) then
        print*, "Stopping due to NaN/Inf in the state at step", k
        exit
      end if
      call write_evol(k, time, (d(1) - d0(1))/d0(1), (d(2) - d0(2))/d0(2), (d(3) - d0(3))/d0(3), &
                      (d(4) - d0(4))/d0(4), (d(5) - d0(5))/d0(5), d(7))
    end do

    call imx_check(imx_swm_get_state(handle, h, u), 'imx_swm_get_state')
    if (dump_state == 1) call dump_final_state()
    call imx_swm_destroy(handle)
    handle = c_null_ptr
  end subroutine swm_tests

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
