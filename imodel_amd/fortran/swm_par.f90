!=======================================================================
! swm_par: the positional parameter files of the reference driver.
!   par/mesh.par  as getparameters  (src/smeshpack.f90:7343-7416), argv
!                 [meshpar] [nv] override as in the reference
!   par/swm.par   as swmpars        (src/swm_data.f90:338-693), the
!                 geo-area subset this build runs: test cases 0, 1, 2, 5,
!                 6, 11, 12, 21, 22, 23, 32-35, 40-42, 51, 52, 53 and the
!                 seeded state 100;
!                 trsk10 and pxt16 with HTC or HC, APVM, diffusion and
!                 hyperdiffusion, gass18 with or without APVM, and the level
!                 model (noPV) with trsk10 and diffusion, on every path;
!                 wrapper none (sinterpol trsk/bary/gass, Coriolis trsk/
!                 pered/gass/dtred, reconmtd trsk/perhx/gass/dubos/melv,
!                 gradmtd trsk) and the level model with other options on
!                 the fp64_ref path
!   par/swm_gpu.par (new, optional): precision, seed, SFC and graph switches
!=======================================================================
module swm_par
  use iso_c_binding, only: c_int32_t
  use datastruct, only: r8
  implicit none
  real(r8), parameter :: day2sec = 86400._r8

  ! mesh.par
  integer :: nv_req = 100, loadable = 0, showonscreen = 0, simulcase = 15, hrchy = 1
  character(len=16) :: mkind = 'icos', mpos = 'eqs', moptm = 'nopt'
  character(len=128) :: mname = ''
  ! swm.par
  integer :: testcase = 2, test_lterror = 0, ntime = 0, adjustntime = 0
  integer :: normal_modes = 0
  integer :: nplots = 0, nprints = 1000000, iploterrors = 0, inopv = 0
  real(r8) :: period = 1._r8, maxtime = 1._r8, dt = 0._r8, gasscoef = 1._r8
  real(r8) :: k2_max = 0._r8, k4_max = 0._r8, hollgw = 2._r8, trskindmax = 0._r8, pvspar = 0.5_r8
  character(len=16) :: stag, mtdwrapper, reconmtd, coriolis_reconmtd, sinterpol, gradmtd
  character(len=16) :: areamtd, diffus, hyperdiffus, pv_stab
  ! swm_gpu.par
  character(len=16) :: precision = 'fp32'
  integer :: seed = 1, use_sfc = 1, use_graph = 1, dump_state = 0

contains

  subroutine read_mesh_par()
    character(len=256) :: filename, buffer, argv
    integer :: fu, nargs, n
    filename = 'par/mesh.par'
    n = 0
    nargs = command_argument_count()
    if (nargs >= 1) call get_command_argument(1, filename)
    if (nargs >= 2) then
      call get_command_argument(2, argv)
      read(argv, *) n
    end if
    open(newunit=fu, file=trim(filename), status='old')
    read(fu, *) buffer
    read(fu, *) buffer
    read(fu, *) nv_req
    read(fu, *) buffer
    read(fu, *) mkind
    read(fu, *) buffer
    read(fu, *) mpos
    read(fu, *) buffer
    read(fu, *) moptm
    read(fu, *) buffer
    read(fu, *) loadable
    read(fu, *) buffer
    read(fu, *) showonscreen
    read(fu, *) buffer
    read(fu, *) simulcase
    read(fu, *) buffer
    read(fu, *) mname
    read(fu, *) buffer
    read(fu, *) hrchy
    close(fu)
    if (n > 0) nv_req = n
  end subroutine read_mesh_par

  ! namegrid (src/smeshpack.f90:7422-7464) for icos/octg grids
  function grid_name(g) result(name)
    integer, intent(in) :: g
    character(len=128) :: name
    character(len=16) :: an, ah
    write(an, '(i8)') g
    if (g == 0) then
      write(ah, '(i2)') 0
    else
      write(ah, '(i2)') hrchy
    end if
    if (trim(moptm) /= 'nopt') then
      name = trim(mkind)//'_'//trim(mpos)//'_'//trim(moptm)//'_h'//trim(adjustl(ah))//'_'//trim(adjustl(an))
    else if (g == 0) then
      name = trim(mkind)//'_'//trim(mpos)//'_'//trim(adjustl(an))
    else
      name = trim(mkind)//'_'//trim(mpos)//'_'//trim(moptm)//'_'//trim(adjustl(an))
    end if
  end function grid_name

  ! nodesizeadj (src/smeshpack.f90:1497-1558): smallest icosahedral level
  ! with at least nv_req nodes.
  integer function glevel_for(nv) result(g)
    integer, intent(in) :: nv
    do g = 0, 11
      if (nv <= 10*4**g + 2) return
    end do
    g = 11
  end function glevel_for

  subroutine read_swm_par(glevel)
    integer, intent(in) :: glevel
    character(len=300) :: buffer
    integer :: fu, ios
    logical :: ex
    open(newunit=fu, file='par/swm.par', status='old')
    read(fu, *) buffer
    read(fu, *) buffer
    read(fu, *) testcase, test_lterror
    read(fu, *) buffer
    read(fu, *) period, maxtime
    read(fu, *) buffer
    read(fu, *) dt, ntime, adjustntime
    read(fu, *) buffer
    read(fu, *) stag
    read(fu, *) buffer
    read(fu, *) mtdwrapper
    read(fu, *) buffer
    read(fu, *) reconmtd, gasscoef
    read(fu, *) buffer
    read(fu, *) coriolis_reconmtd
    read(fu, *) buffer
    read(fu, *) sinterpol
    read(fu, *) buffer
    read(fu, *) gradmtd
    read(fu, *) buffer
    read(fu, *) areamtd
    read(fu, *) buffer
    read(fu, *) nplots, nprints, iploterrors
    read(fu, *) buffer
    read(fu, *) k2_max, diffus
    read(fu, *) buffer
    read(fu, *) k4_max, hyperdiffus
    read(fu, *) buffer
    read(fu, *) hollgw
    read(fu, *) buffer
    read(fu, *) trskindmax
    read(fu, *) buffer
    read(fu, *) pv_stab, pvspar
    read(fu, *) buffer
    read(fu, *) inopv
    close(fu)

    ! the scheme this build runs (the reference defaults, par/swm.par)
    if (trim(stag) /= 'HTC' .and. trim(stag) /= 'HC') call unsupported('staggering', stag)
    if (wrap_code() < 0) then
      if (sint_code() < 0) call unsupported('scalar interpolation', sinterpol)
      if (cor_code() < 0) call unsupported('Coriolis reconstruction', coriolis_reconmtd)
      if (rec_code() < 0) call unsupported('vector reconstruction', reconmtd)
      call unsupported('gradient method', gradmtd)
    end if
    if (inopv > 0 .and. (cor_code() == 1 .or. cor_code() == 3)) then
      ! as swmpars (src/swm_data.f90:490-496)
      print*, "Cannot use this Coriolis vector reconstruction method with level model (noPV)", &
        coriolis_reconmtd, .true.
      stop 1
    end if
    if (sint_code() == 2 .and. (k2_max > 0._r8 .or. k4_max > 0._r8)) then
      ! as the reference's scalar_hx2tr (src/swm_operators.f90:271-273)
      print*, "Error at scalar_hx2trcc: Gassmann's scheme needs edge values"
      stop 1
    end if
    if (areamtd(1:3) /= 'geo') call unsupported('area method', areamtd)
    if (k2_max > 0._r8) then
      if (coef_code(diffus) < 0) call unsupported('diffusion coefficient function', diffus)
    end if
    if (k4_max > 0._r8) then
      if (coef_code(hyperdiffus) < 0) call unsupported('hyperdiffusion coefficient function', hyperdiffus)
    end if
    if (trim(pv_stab) /= 'none' .and. trim(pv_stab) /= 'apvm') call unsupported('PV stabilisation', pv_stab)
    if (test_lterror /= 0) call unsupported('local truncation error test', '1')
    select case (testcase)
    case (0, 1, 2, 5, 6, 11, 12, 21, 22, 23, 32, 33, 34, 35, 40, 41, 42, 51, 52, 53, 100)
    case default
      call unsupported('test case (0, 1, 2, 5, 6, 11, 12, 21, 22, 23, 32-35, 40-42, 51, 52, 53 or the seeded 100)', &
                       'other')
    end select

    ! dt / ntime (src/swm_data.f90:562-586)
    if (period <= 0) stop "swmpars error: period should be positive and must be given in days"
    period = period*day2sec
    maxtime = maxtime*day2sec
    if (dt > 0) then
      if (adjustntime == 1) dt = dt*(2**(5 - glevel))
      ntime = ceiling(maxtime/dt)
    else if (dt == 0 .and. ntime > 0) then
      if (adjustntime == 1) ntime = ntime*(2**(glevel - 5))
      dt = maxtime/real(ntime, r8)
    else
      dt = 50
      if (adjustntime == 1) dt = dt*(2**(5 - glevel))
      ntime = ceiling(maxtime/dt)
    end if
    if (nprints <= 0) nprints = 1000000

    inquire(file='par/swm_gpu.par', exist=ex)
    if (ex) then
      open(newunit=fu, file='par/swm_gpu.par', status='old')
      read(fu, *, iostat=ios) buffer
      read(fu, *, iostat=ios) precision
      read(fu, *, iostat=ios) buffer
      read(fu, *, iostat=ios) seed
      read(fu, *, iostat=ios) buffer
      read(fu, *, iostat=ios) use_sfc, use_graph
      read(fu, *, iostat=ios) buffer
      read(fu, *, iostat=ios) dump_state
      read(fu, *, iostat=ios) buffer
      read(fu, *, iostat=ios) normal_modes
      if (ios /= 0) normal_modes = 0
      close(fu)
    end if
  end subroutine read_swm_par

  integer(c_int32_t) function precision_code() result(p)
    select case (trim(precision))
    case ('fp64_ref'); p = 0
    case ('fp64'); p = 1
    case ('fp32'); p = 2
    case default
      call unsupported('precision', precision)
      p = 2
    end select
  end function precision_code

  ! diffusion coefficient function -> IMX_COEF_* (-1: unknown)
  integer(c_int32_t) function coef_code(name) result(c)
    character(len=*), intent(in) :: name
    select case (trim(name))
    case ('const'); c = 0
    case ('align'); c = 1
    case ('diam'); c = 2
    case default; c = -1
    end select
  end function coef_code

  ! method wrapper -> IMX_WRAP_* (-1: not part of this build). The
  ! explicit form "none" with trsk in the four method lines is trsk10; any
  ! other known triple with the trsk gradient is IMX_WRAP_NONE.
  integer(c_int32_t) function wrap_code() result(c)
    select case (trim(mtdwrapper))
    case ('trsk10'); c = 0
    case ('pxt16'); c = 1
    case ('gass18'); c = 2
    case ('none')
      c = -1
      if (sint_code() >= 0 .and. cor_code() >= 0 .and. rec_code() >= 0 .and. trim(gradmtd) == 'trsk') c = 3
      if (c == 3 .and. sint_code() == 0 .and. cor_code() == 0 .and. rec_code() == 0) c = 0
    case default; c = -1
    end select
  end function wrap_code

  ! the scalar interpolation, Coriolis and kinetic-energy methods the
  ! wrapper stands for, or the named ones with wrapper none
  ! (src/swm_data.f90:445-522) -> IMX_SINT_*, IMX_COR_*, IMX_REC_* (-1:
  ! not part of this build: hyb and exact Coriolis)
  integer(c_int32_t) function sint_code() result(c)
    select case (trim(mtdwrapper))
    case ('trsk10'); c = 0
    case ('pxt16'); c = 1
    case ('gass18'); c = 2
    case default
      select case (trim(sinterpol))
      case ('trsk'); c = 0
      case ('bary'); c = 1
      case ('gass'); c = 2
      case default; c = -1
      end select
    end select
  end function sint_code

  integer(c_int32_t) function cor_code() result(c)
    select case (trim(mtdwrapper))
    case ('trsk10'); c = 0
    case ('pxt16'); c = 1
    case ('gass18'); c = 2
    case default
      select case (trim(coriolis_reconmtd))
      case ('trsk'); c = 0
      case ('pered'); c = 1
      case ('gass'); c = 2
      case ('dtred'); c = 3
      case default; c = -1
      end select
    end select
  end function cor_code

  integer(c_int32_t) function rec_code() result(c)
    select case (trim(mtdwrapper))
    case ('trsk10', 'gass18'); c = 0
    case ('pxt16'); c = 1
    case default
      select case (trim(reconmtd))
      case ('trsk'); c = 0
      case ('perhx'); c = 1
      case ('gass', 'dubos'); c = 2
      case ('melv'); c = 3
      case default; c = -1
      end select
    end select
  end function rec_code

  ! gasscoef of the gass kinetic energy; dubos is gass with 0
  real(r8) function gass_coef() result(g)
    g = gasscoef
    if (trim(mtdwrapper) == 'none' .and. trim(reconmtd) == 'dubos') g = 0._r8
  end function gass_coef

  ! a wrapper other than trsk10, APVM, diffusion, hyperdiffusion, HC or the
  ! level model selected
  logical function scheme_options() result(on)
    on = wrap_code() /= 0 .or. trim(pv_stab) == 'apvm' .or. k2_max > 0._r8 .or. k4_max > 0._r8 .or. &
         trim(stag) == 'HC' .or. inopv > 0
  end function scheme_options

  subroutine unsupported(what, val)
    character(len=*), intent(in) :: what, val
    print*, "ERROR: ", what, " '", trim(val), "' is not part of this build"
    stop 1
  end subroutine unsupported
end module swm_par
