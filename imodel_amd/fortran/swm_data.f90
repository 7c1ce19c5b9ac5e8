!=======================================================================
! swm_data: the module globals of the shallow-water model that callers of
! the swm module read and write (src/swm_data.f90:31-334): the mesh, the
! state h and u, the bottom topography bt, the equation right-hand sides
! masseq and momeq, the exact fields of the test cases with an analytic
! solution, and the parameters of par/swm.par (module swm_par).
!   swmpars(usetime)           reads par/swm.par (:338-693)
!   allocate_globalswmvars()   sizes the fields on the mesh (:695-1084)
!=======================================================================
module swm_data
  use iso_c_binding
  use datastruct
  use swm_par
  implicit none
  type(grid_structure) :: mesh
  type(scalar_field) :: h, u, bt, masseq, momeq
  type(scalar_field) :: h_exact, u_exact
  ! the handle the fields are evolved on (imx_swm), opened by
  ! initialize_gridprop
  type(c_ptr) :: handle = c_null_ptr
  logical :: usetime_names = .true.
  ! the Coriolis term of the test case (src/swm_data.f90:42-43): 0 variable
  ! f, 1 linearised (eta = fcte), 2 f-sphere (f = fcte)
  integer :: fsphere = 0
  real(r8) :: fcte = 0._r8

contains

  subroutine swmpars(usetime)
    logical, intent(in) :: usetime
    usetime_names = usetime
    call read_swm_par(mesh%glevel)
  end subroutine swmpars

  subroutine allocate_globalswmvars()
    integer :: edpos
    edpos = 6                              ! HTC: edges at the tr x hx crossings
    if (trim(stag) == 'HC') edpos = 3      ! HC: Voronoi edge midpoints
    call field(h, mesh%nv, 0, 'h')
    call field(bt, mesh%nv, 0, 'bt')
    call field(masseq, mesh%nv, 0, 'masseq')
    call field(h_exact, mesh%nv, 0, 'h_exact')
    call field(u, mesh%ne, edpos, 'u')
    call field(momeq, mesh%ne, edpos, 'momeq')
    call field(u_exact, mesh%ne, edpos, 'u_exact')
  contains
    subroutine field(f, n, pos, name)
      type(scalar_field), intent(inout) :: f
      integer, intent(in) :: n, pos
      character(len=*), intent(in) :: name
      if (allocated(f%f)) deallocate(f%f)
      allocate(f%f(n))
      f%f = 0._r8
      f%n = n
      f%pos = pos
      f%name = name
    end subroutine field
  end subroutine allocate_globalswmvars
end module swm_data
