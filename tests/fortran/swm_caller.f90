!=======================================================================
! TEST INFRASTRUCTURE: a caller written against the reference's swm call
! surface, in the shape of moist_swm (src/moist_swm.f90:55-66, :1205):
! it imports the mesh and field types, the swm_data module globals and
! `initialize_gridprop, sumf_areas, calc_energies, tendency, error_calc`
! (plus ode_rk4) from swm, and drives them as the reference's routines are
! driven. Reads par/mesh.par and par/swm.par from the working directory and
! writes out.bin (IMXBIN records): masseq, momeq of the initial state, the
! initial mass and energies, and h, u after 3 RK4 steps.
!=======================================================================
program swm_caller
  use datastruct, only: grid_structure, scalar_field, r8
  use smeshpack, only: getparameters, meshbuild
  use swm_data
  use swm, only: initialize_gridprop, initialize_fields, sumf_areas, calc_energies, tendency, &
                 error_calc, ode_rk4
  implicit none
  type(grid_structure) :: meshtmp
  type(scalar_field) :: h_new, u_new, h_err
  real(r8) :: tmass, pe, ke, te, ae, emr, e2, em, time
  integer :: k, fu

  call getparameters(meshtmp)
  call meshbuild(meshtmp)
  mesh = meshtmp
  call swmpars(usetime=.true.)
  call allocate_globalswmvars()
  call initialize_gridprop()
  call initialize_fields()
  if (testcase > 1) call tendency(h, u, masseq%f, momeq%f)
  tmass = sumf_areas(h)
  call calc_energies(pe, ke, te, ae)
  h_new = h
  u_new = u
  do k = 1, 3
    time = real(k, r8)*dt
    call ode_rk4(time, h, u, h_new, u_new, dt)
    h = h_new
    u = u_new
  end do
  call error_calc(h, h_exact, h_err, emr, e2, em)

  open(newunit=fu, file='out.bin', access='stream', form='unformatted', status='replace')
  write(fu) 'IMXBIN01'
  call rec('masseq', masseq%f)
  call rec('momeq', momeq%f)
  call rec('diag0', [tmass, pe, ke, te, ae])
  call rec('h', h%f)
  call rec('u', u%f)
  call rec('herr', [emr, e2, em])
  close(fu)
contains
  subroutine rec(name, a)
    character(len=*), intent(in) :: name
    real(r8), intent(in) :: a(:)
    character(len=24) :: nm
    nm = name
    write(fu) nm, 'f8  ', int(size(a), 8), a
  end subroutine rec
end program swm_caller
