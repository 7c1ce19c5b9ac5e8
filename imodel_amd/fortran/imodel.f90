!=======================================================================
! imodel: driver of the MI355X shallow-water build (src/imodel.f90:1-170).
! Reads par/mesh.par into the mesh structure (getparameters), gets the
! grid as the reference's meshbuild does (module smeshpack: loaded from
! grid/ when loadable = 1 and the files exist, otherwise built with the
! reference's triangulation order and, with loadable = 1, stored), then runs
! simulcase 15, the shallow-water model, or simulcase 14, its local
! truncation error analysis (src/imodel.f90:149-153), with the mesh as
! their argument as in the reference.
!=======================================================================
program imodel
  use datastruct, only: grid_structure
  use smeshpack, only: getparameters, meshbuild
  use swm_par, only: simulcase
  use swm, only: swm_tests, swm_horiz_loc_trunc_er
  use imx_c, only: imx_mesh_destroy
  implicit none
  type(grid_structure) :: mesh

  call getparameters(mesh)
  call meshbuild(mesh)

  select case (simulcase)
  case (14)
    call swm_horiz_loc_trunc_er(mesh)
  case (15)
    call swm_tests(mesh)
  case default
    print*, "simulcase", simulcase, "is not part of this build (14: local truncation error, ", &
      "15: shallow water model)"
    stop 1
  end select
  call imx_mesh_destroy(mesh%lib)
end program imodel
