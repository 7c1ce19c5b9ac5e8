!=======================================================================
! imodel: driver of the MI355X shallow-water build (src/imodel.f90:1-170).
! Reads par/mesh.par, builds the icosahedral non-optimised grid with
! libimx (meshbuild for icos/nopt, src/smeshpack.f90:77-368) and runs
! simulcase 15, the shallow-water model (src/imodel.f90:152-153).
!=======================================================================
program imodel
  use iso_c_binding
  use imx_c
  use swm_par
  use swm, only: swm_tests
  implicit none
  type(c_ptr) :: meshp = c_null_ptr
  integer :: g
  integer(c_int32_t) :: pos

  call read_mesh_par()
  if (trim(mkind) /= 'icos') call unsupported('mesh kind', mkind)
  if (trim(moptm) /= 'nopt') call unsupported('mesh optimisation', moptm)
  select case (trim(mpos))
  case ('eqs'); pos = IMX_POS_EQS
  case ('pol'); pos = IMX_POS_POL
  case default; call unsupported('mesh position', mpos); pos = 0
  end select
  g = glevel_for(nv_req)
  print*, "Mesh: icos ", trim(mpos), " nopt, glevel", g, " nodes", 10*4**g + 2
  call imx_check(imx_mesh_build_icos(int(g, c_int32_t), pos, meshp), 'meshbuild')
  call imx_check(imx_mesh_gridprop(meshp), 'initialize_gridprop')

  select case (simulcase)
  case (15)
    call swm_tests(meshp, g)
  case default
    print*, "simulcase", simulcase, "is not part of this build (15: shallow water model)"
    stop 1
  end select
  call imx_mesh_destroy(meshp)
end program imodel
