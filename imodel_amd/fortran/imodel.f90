!=======================================================================
! imodel: driver of the MI355X shallow-water build (src/imodel.f90:1-170).
! Reads par/mesh.par and gets the grid as the reference's meshbuild does
! (src/smeshpack.f90:77-368): with loadable = 1 and grid/<name>_header.dat
! present, the grid files are loaded (meshload layout, so grids the
! reference built or optimised, e.g. SCVT, run unchanged); otherwise the
! icosahedral non-optimised grid is built with libimx and, with
! loadable = 1, stored in grid/ for later runs (meshstore). Then runs
! simulcase 15, the shallow-water model, or simulcase 14, its local
! truncation error analysis (src/imodel.f90:149-153).
!=======================================================================
program imodel
  use iso_c_binding
  use imx_c
  use swm_par
  use swm, only: swm_tests, swm_horiz_loc_trunc_er
  implicit none
  type(c_ptr) :: meshp = c_null_ptr
  integer :: g
  integer(c_int32_t) :: pos
  character(len=128) :: gname
  logical :: have

  call read_mesh_par()
  g = glevel_for(nv_req)
  gname = grid_name(g)
  if (trim(mkind) == 'read') then
    ! namegrid for read grids: the node file's name without its extension
    gname = mname(1:len_trim(mname)-4)
    if (trim(moptm) /= 'nopt') gname = trim(gname)//'_'//trim(moptm)
    g = -1
  end if
  gridname = gname
  inquire(file='grid/'//trim(gname)//'_header.dat', exist=have)
  if (loadable == 1 .and. have) then
    print*, "Mesh: loading grid/", trim(gname)
    call imx_check(imx_mesh_load_ref('grid/'//c_null_char, trim(gname)//c_null_char, meshp), 'meshload')
  else if (trim(mkind) == 'read') then
    if (trim(moptm) /= 'nopt') call unsupported('mesh optimisation (without its grid files)', moptm)
    print*, "Mesh: reading grid/", trim(mname)
    call imx_check(imx_mesh_build_read('grid/'//trim(mname)//c_null_char, meshp), 'meshread')
    if (loadable == 1) then
      call imx_check(imx_mesh_write_ref(meshp, 'grid/'//c_null_char, trim(gname)//c_null_char), 'meshstore')
    end if
  else
    if (trim(mkind) /= 'icos') call unsupported('mesh kind', mkind)
    if (trim(moptm) /= 'nopt') call unsupported('mesh optimisation (without its grid files)', moptm)
    select case (trim(mpos))
    case ('eqs'); pos = IMX_POS_EQS
    case ('pol'); pos = IMX_POS_POL
    case default; call unsupported('mesh position', mpos); pos = 0
    end select
    print*, "Mesh: icos ", trim(mpos), " nopt, glevel", g, " nodes", 10*4**g + 2
    call imx_check(imx_mesh_build_icos(int(g, c_int32_t), pos, meshp), 'meshbuild')
    if (loadable == 1) then
      call imx_check(imx_mesh_write_ref(meshp, 'grid/'//c_null_char, trim(gname)//c_null_char), 'meshstore')
    end if
  end if
  call imx_check(imx_mesh_gridprop(meshp), 'initialize_gridprop')

  select case (simulcase)
  case (14)
    call swm_horiz_loc_trunc_er(meshp, g)
  case (15)
    call swm_tests(meshp, g)
  case default
    print*, "simulcase", simulcase, "is not part of this build (14: local truncation error, ", &
      "15: shallow water model)"
    stop 1
  end select
  call imx_mesh_destroy(meshp)
end program imodel
