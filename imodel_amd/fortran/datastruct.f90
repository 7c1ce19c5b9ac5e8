!=======================================================================
! datastruct: the derived types the reference's swm call surface passes
! (src/datastruct.f90:403-589), for callers written against it.
!   scalar_field   f(:), n, pos, name: the same components as the
!                  reference's (pos 0 cells, 1 triangles, 6 edges)
!   grid_structure the mesh parameters of the reference's structure (nv,
!                  ne, nt, glevel, kind, pos, optm, loadable, hrchy, name);
!                  its tables live in the library mesh `lib` (imx_mesh)
!                  instead of allocatable per-entity components, so the
!                  structure is copied by value as the reference's is
!                  (mesh = meshtmp, src/swm.f90:152)
!=======================================================================
module datastruct
  use iso_c_binding, only: c_ptr, c_null_ptr
  implicit none
  integer, parameter :: i4 = selected_int_kind(6)
  integer, parameter :: r8 = selected_real_kind(12, 100)

  type scalar_field
    real(r8), allocatable :: f(:)
    integer(i4) :: n = 0
    integer(i4) :: pos = 0
    character(len=256) :: name = ''
  end type scalar_field

  type grid_structure
    integer(i4) :: nv = 0, ne = 0, nt = 0
    integer(i4) :: glevel = -1
    integer(i4) :: maxvnb = 6
    integer(i4) :: loadable = 0, hrchy = 0
    character(len=128) :: kind = 'icos', pos = 'eqs', optm = 'nopt'
    character(len=128) :: name = '', filename = ''
    type(c_ptr) :: lib = c_null_ptr
  end type grid_structure
end module datastruct
