!=======================================================================
! smeshpack: the mesh entry points of the reference's driver
! (src/smeshpack.f90) over libimx.
!   getparameters(mesh)  par/mesh.par into the structure (:7343-7416)
!   namegrid(mesh)       the grid name of the file set (:7422-7464)
!   meshbuild(mesh)      icos/nopt or read grids, loaded from grid/ when
!                        loadable = 1 and the files exist, else built and,
!                        with loadable = 1, stored (:77-368); the library
!                        builds with the reference's triangulation order,
!                        so every table carries the reference's numbering
!=======================================================================
module smeshpack
  use iso_c_binding
  use datastruct
  use imx_c
  use swm_par, only: read_mesh_par, nv_req, mkind, mpos, moptm, loadable, hrchy, mname, glevel_for
  implicit none

contains

  subroutine getparameters(mesh)
    type(grid_structure), intent(inout) :: mesh
    call read_mesh_par()
    mesh%nv = nv_req
    mesh%kind = mkind
    mesh%pos = mpos
    mesh%optm = moptm
    mesh%loadable = loadable
    mesh%hrchy = hrchy
    mesh%filename = mname
  end subroutine getparameters

  subroutine namegrid(mesh)
    type(grid_structure), intent(inout) :: mesh
    character(len=16) :: an, ah
    integer :: n
    if (trim(mesh%kind) == 'read') then
      n = len_trim(mesh%filename) - 4
      mesh%name = mesh%filename(1:n)
      if (trim(mesh%optm) /= 'nopt') mesh%name = trim(mesh%name)//'_'//trim(mesh%optm)
      return
    end if
    write(an, '(i8)') mesh%glevel
    if (mesh%glevel == 0) then
      write(ah, '(i2)') 0
    else
      write(ah, '(i2)') mesh%hrchy
    end if
    if (trim(mesh%optm) /= 'nopt') then
      mesh%name = trim(mesh%kind)//'_'//trim(mesh%pos)//'_'//trim(mesh%optm)//'_h'//trim(adjustl(ah))//'_'// &
                  trim(adjustl(an))
    else if (mesh%glevel == 0) then
      mesh%name = trim(mesh%kind)//'_'//trim(mesh%pos)//'_'//trim(adjustl(an))
    else
      mesh%name = trim(mesh%kind)//'_'//trim(mesh%pos)//'_'//trim(mesh%optm)//'_'//trim(adjustl(an))
    end if
  end subroutine namegrid

  subroutine meshbuild(mesh)
    type(grid_structure), intent(inout) :: mesh
    type(imx_mesh_desc) :: d
    integer(c_int32_t) :: pos
    logical :: have
    if (trim(mesh%kind) == 'read') then
      mesh%glevel = -1
    else
      mesh%glevel = glevel_for(mesh%nv)
      mesh%nv = 10*4**mesh%glevel + 2
    end if
    call namegrid(mesh)
    inquire(file='grid/'//trim(mesh%name)//'_header.dat', exist=have)
    if (mesh%loadable == 1 .and. have) then
      print*, "Mesh: loading grid/", trim(mesh%name)
      call imx_check(imx_mesh_load_ref('grid/'//c_null_char, trim(mesh%name)//c_null_char, mesh%lib), 'meshload')
    else if (trim(mesh%kind) == 'read') then
      if (trim(mesh%optm) /= 'nopt') call unsupported_mesh('mesh optimisation (without its grid files)', mesh%optm)
      print*, "Mesh: reading grid/", trim(mesh%filename)
      call imx_check(imx_mesh_build_read('grid/'//trim(mesh%filename)//c_null_char, mesh%lib), 'meshread')
      if (mesh%loadable == 1) call store()
    else
      if (trim(mesh%kind) /= 'icos') call unsupported_mesh('mesh kind', mesh%kind)
      if (trim(mesh%optm) /= 'nopt') call unsupported_mesh('mesh optimisation (without its grid files)', mesh%optm)
      select case (trim(mesh%pos))
      case ('eqs'); pos = IMX_POS_EQS
      case ('pol'); pos = IMX_POS_POL
      case default; call unsupported_mesh('mesh position', mesh%pos); pos = 0
      end select
      print*, "Mesh: icos ", trim(mesh%pos), " nopt, glevel", mesh%glevel, " nodes", mesh%nv
      call imx_check(imx_mesh_build_icos(int(mesh%glevel, c_int32_t), pos, mesh%lib), 'meshbuild')
      if (mesh%loadable == 1) call store()
    end if
    call imx_check(imx_mesh_desc_of(mesh%lib, d), 'meshbuild')
    mesh%nv = d%nv
    mesh%ne = d%ne
    mesh%nt = d%nt
  contains
    subroutine store()
      call imx_check(imx_mesh_write_ref(mesh%lib, 'grid/'//c_null_char, trim(mesh%name)//c_null_char), 'meshstore')
    end subroutine store
  end subroutine meshbuild

  subroutine unsupported_mesh(what, val)
    character(len=*), intent(in) :: what, val
    print*, "ERROR: ", what, " '", trim(val), "' is not part of this build"
    stop 1
  end subroutine unsupported_mesh
end module smeshpack
