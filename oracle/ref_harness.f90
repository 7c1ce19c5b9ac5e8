!=======================================================================
! ref_harness -- TEST INFRASTRUCTURE ONLY (never shipped, never timed as
! the product). A driver of our own that links the reference iModel SWM
! modules, compiled from their sources in place by oracle/Makefile.ref,
! and writes golden vectors / reference timings in the IMXBIN container
! read by tests/ and bench.py (cpu_baseline leg).
!
! It calls, unchanged, the reference routines on the SWM hot path:
!   meshbuild        src/smeshpack.f90:77-368
!   swmpars          src/swm_data.f90:338-693
!   allocate_globalswmvars / initialize_gridprop / initialize_fields
!                    src/swm_data.f90:695, src/swm.f90:979-1041, :1043
!   tendency         src/swm.f90:2416-2598
!   ode_rk4          src/swm.f90:2600-2683
!   sumf_areas, sumfsq_areas, calc_energies   src/swm.f90:3394-3490, :2686
! and mirrors the time loop of swm_tests (src/swm.f90:238-370).
!
! Usage (run in a scratch directory; par/ grid/ data/ are created there):
!   ref_harness build <glevel> <eqs|pol> <out.bin>
!   ref_harness run   <mesh: build:G:pos | file.bin> <state: tc2 | file.bin>
!                     <nsteps> <dt> <diag 0|1> <out.bin> [scheme]
!   ref_harness buildread <name.xyz|name.gmt> <nv> <out.bin>
!                     meshbuild for mesh%kind = 'read' from grid/<name> and
!                     grid/<name without extension>.ngb, then as build
!   ref_harness lte   <mesh: build:G:pos | file.bin> <out.bin> [scheme]
!                     swm_horiz_loc_trunc_er (src/swm.f90:382-734) on TC2:
!                     its table and mimetic sums go to stdout; out.bin gets
!                     error_calc / error_calc_vec of every row it prints
!                     (lte_<row> = maxrel, rel L2, max) and exact fields
! scheme: comma-separated key=value lines of par/swm.par, e.g.
!   "stag=HTC,wrap=trsk10,k2=1e5,diffus=const,k4=0,hyperdiffus=const,pv=apvm,pvspar=0.5"
! (defaults: HTC, trsk10, no diffusion, no PV stabilisation); the
! diffusion coefficients are then set up as swm_tests does
! (src/swm.f90:170-210).
!=======================================================================
module imxbin
  use constants, only: i4, r8
  implicit none
  integer :: ounit
  type rec_t
    character(len=24) :: name
    character(len=4)  :: dt
    integer(i4), allocatable :: iv(:)
    real(r8), allocatable :: rv(:)
  end type rec_t
  type(rec_t), allocatable :: recs(:)
  integer :: nrec = 0
contains
  subroutine bin_open_w(fname)
    character(len=*), intent(in) :: fname
    open(newunit=ounit, file=fname, access='stream', form='unformatted', status='replace')
    write(ounit) 'IMXBIN01'
  end subroutine bin_open_w

  subroutine put_i4(name, a)
    character(len=*), intent(in) :: name
    integer(i4), intent(in) :: a(:)
    character(len=24) :: nm
    nm = name
    write(ounit) nm, 'i4  ', int(size(a), 8), a
  end subroutine put_i4

  subroutine put_f8(name, a)
    character(len=*), intent(in) :: name
    real(r8), intent(in) :: a(:)
    character(len=24) :: nm
    nm = name
    write(ounit) nm, 'f8  ', int(size(a), 8), a
  end subroutine put_f8

  subroutine bin_close()
    close(ounit)
  end subroutine bin_close

  subroutine bin_read(fname)
    character(len=*), intent(in) :: fname
    integer :: iu, ios
    character(len=8) :: magic
    character(len=24) :: nm
    character(len=4) :: dt
    integer(8) :: cnt
    type(rec_t), allocatable :: tmp(:)
    if (allocated(recs)) deallocate(recs)
    allocate(recs(256))
    nrec = 0
    open(newunit=iu, file=fname, access='stream', form='unformatted', status='old')
    read(iu) magic
    if (magic /= 'IMXBIN01') stop 'bin_read: bad magic'
    do
      read(iu, iostat=ios) nm, dt, cnt
      if (ios /= 0) exit
      nrec = nrec + 1
      if (nrec > size(recs)) then
        allocate(tmp(2*size(recs)))
        tmp(1:size(recs)) = recs
        call move_alloc(tmp, recs)
      end if
      recs(nrec)%name = nm
      recs(nrec)%dt = dt
      if (dt == 'i4  ') then
        allocate(recs(nrec)%iv(cnt))
        read(iu) recs(nrec)%iv
      else
        allocate(recs(nrec)%rv(cnt))
        read(iu) recs(nrec)%rv
      end if
    end do
    close(iu)
  end subroutine bin_read

  function find_rec(name) result(k)
    character(len=*), intent(in) :: name
    integer :: k
    do k = 1, nrec
      if (trim(recs(k)%name) == trim(name)) return
    end do
    print*, 'bin_read: missing record ', trim(name)
    stop 2
  end function find_rec

  function geti(name) result(a)
    character(len=*), intent(in) :: name
    integer(i4), allocatable :: a(:)
    integer :: k
    k = find_rec(name)
    a = recs(k)%iv
  end function geti

  function getr(name) result(a)
    character(len=*), intent(in) :: name
    real(r8), allocatable :: a(:)
    integer :: k
    k = find_rec(name)
    a = recs(k)%rv
  end function getr
end module imxbin

module harness_mesh
  use constants, only: i4, r8
  use datastruct, only: grid_structure
  use imxbin
  implicit none
  integer, parameter :: mx = 6
contains
  !Flatten the reference grid_structure into the IMXBIN mesh records
  !(reference numbering, 1-based; per-cell lists padded to 6 with 0).
  subroutine dump_mesh(mesh, with_gridprop)
    type(grid_structure), intent(in) :: mesh
    logical, intent(in) :: with_gridprop
    integer(i4), allocatable :: ia(:)
    real(r8), allocatable :: ra(:)
    integer :: i, j, l, k, n, nv, ne, nt

    nv = mesh%nv; ne = mesh%ne; nt = mesh%nt
    call put_i4('dims', (/ nv, ne, nt, mesh%glevel /))
    allocate(ra(3*max(nv, ne, nt)))
    do i = 1, nv
      ra(3*i-2:3*i) = mesh%v(i)%p
    end do
    call put_f8('v_p', ra(1:3*nv))
    call put_f8('v_lat', (/ (mesh%v(i)%lat, i=1,nv) /))
    call put_f8('v_lon', (/ (mesh%v(i)%lon, i=1,nv) /))
    call put_i4('v_nnb', (/ (mesh%v(i)%nnb, i=1,nv) /))
    allocate(ia(mx*max(nv, 1)))
    call pack_cell_i(1); call put_i4('v_nb', ia(1:mx*nv))
    call pack_cell_i(2); call put_i4('v_ed', ia(1:mx*nv))
    call pack_cell_i(3); call put_i4('v_tr', ia(1:mx*nv))
    call pack_cell_i(4); call put_i4('hx_ttgout', ia(1:mx*nv))
    call pack_cell_i(5); call put_i4('hx_tnrccw', ia(1:mx*nv))
    call pack_cell_i(6); call put_i4('hx_tg', ia(1:mx*nv))
    call pack_cell_i(7); call put_i4('hx_nr', ia(1:mx*nv))
    call put_f8('hx_areag', (/ (mesh%hx(i)%areag, i=1,nv) /))

    deallocate(ia); allocate(ia(3*max(ne, nt)))
    do l = 1, ne
      ia(2*l-1:2*l) = mesh%ed(l)%v
    end do
    call put_i4('ed_v', ia(1:2*ne))
    do l = 1, ne
      ia(2*l-1:2*l) = mesh%ed(l)%sh
    end do
    call put_i4('ed_sh', ia(1:2*ne))
    call put_f8('ed_leng', (/ (mesh%ed(l)%leng, l=1,ne) /))
    call put_f8('edhx_leng', (/ (mesh%edhx(l)%leng, l=1,ne) /))
    do l = 1, ne
      ra(3*l-2:3*l) = mesh%ed(l)%c%p
    end do
    call put_f8('ed_c', ra(1:3*ne))
    call put_f8('ed_clat', (/ (mesh%ed(l)%c%lat, l=1,ne) /))
    call put_f8('ed_clon', (/ (mesh%ed(l)%c%lon, l=1,ne) /))
    do l = 1, ne
      ra(3*l-2:3*l) = mesh%ed(l)%tg
    end do
    call put_f8('ed_tg', ra(1:3*ne))
    do l = 1, ne
      ra(3*l-2:3*l) = mesh%ed(l)%nr
    end do
    call put_f8('ed_nr', ra(1:3*ne))
    do l = 1, ne
      ra(3*l-2:3*l) = mesh%edhx(l)%c%p
    end do
    call put_f8('edhx_c', ra(1:3*ne))
    do l = 1, ne
      ra(3*l-2:3*l) = mesh%edhx(l)%nr
    end do
    call put_f8('edhx_nr', ra(1:3*ne))
    do l = 1, ne
      ra(3*l-2:3*l) = mesh%edhx(l)%tg
    end do
    call put_f8('edhx_tg', ra(1:3*ne))

    do k = 1, nt
      ia(3*k-2:3*k) = mesh%tr(k)%v
    end do
    call put_i4('tr_v', ia(1:3*nt))
    do k = 1, nt
      ia(3*k-2:3*k) = mesh%tr(k)%ed
    end do
    call put_i4('tr_ed', ia(1:3*nt))
    do k = 1, nt
      ia(3*k-2:3*k) = mesh%tr(k)%tg
    end do
    call put_i4('tr_tg', ia(1:3*nt))
    do k = 1, nt
      ia(3*k-2:3*k) = mesh%tr(k)%nr
    end do
    call put_i4('tr_nr', ia(1:3*nt))
    call put_f8('tr_areag', (/ (mesh%tr(k)%areag, k=1,nt) /))
    do k = 1, nt
      ra(3*k-2:3*k) = mesh%tr(k)%c%p
    end do
    call put_f8('tr_c', ra(1:3*nt))
    call put_f8('tr_clat', (/ (mesh%tr(k)%c%lat, k=1,nt) /))
    call put_f8('tr_clon', (/ (mesh%tr(k)%c%lon, k=1,nt) /))

    if (with_gridprop) then
      do k = 1, nt
        ra(3*k-2:3*k) = mesh%tr(k)%trhx_area
      end do
      call put_f8('tr_trhx_area', ra(1:3*nt))
      deallocate(ra); allocate(ra(mx*mx*nv))
      ra = 0._r8
      do i = 1, nv
        n = mesh%v(i)%nnb
        ra(mx*(i-1)+1:mx*(i-1)+n) = mesh%hx(i)%hxtr_area(1:n)
      end do
      call put_f8('hx_hxtr_area', ra(1:mx*nv))
      ra = 0._r8
      do i = 1, nv
        n = mesh%v(i)%nnb
        do j = 1, n
          do l = 1, n
            !row-major: [cell][i][j] = trskw(i, j)
            ra(mx*mx*(i-1) + mx*(j-1) + l) = mesh%hx(i)%trskw(j, l)
          end do
        end do
      end do
      call put_f8('hx_trskw', ra(1:mx*mx*nv))
    end if
  contains
    subroutine pack_cell_i(which)
      integer, intent(in) :: which
      integer :: ii, nn
      ia = 0
      do ii = 1, nv
        nn = mesh%v(ii)%nnb
        if (nn > mx) stop 'dump_mesh: more than 6 neighbours'
        select case (which)
        case (1); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%v(ii)%nb(1:nn)
        case (2); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%v(ii)%ed(1:nn)
        case (3); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%v(ii)%tr(1:nn)
        case (4); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%hx(ii)%ttgout(1:nn)
        case (5); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%hx(ii)%tnrccw(1:nn)
        case (6); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%hx(ii)%tg(1:nn)
        case (7); ia(mx*(ii-1)+1:mx*(ii-1)+nn) = mesh%hx(ii)%nr(1:nn)
        end select
      end do
    end subroutine pack_cell_i
  end subroutine dump_mesh

  !Fill the reference grid_structure from IMXBIN mesh records (every
  !component the trsk10/HTC tendency, gridprop, TC2 set-up and the
  !diagnostics read).
  subroutine load_mesh(fname, mesh)
    character(len=*), intent(in) :: fname
    type(grid_structure), intent(inout) :: mesh
    integer(i4), allocatable :: dims(:), nnb(:), nb(:), ved(:), vtr(:), ttg(:), tnr(:), htg(:), hnr(:)
    integer(i4), allocatable :: edv(:), edsh(:), trv(:), tred(:), trtg(:), trnr(:)
    real(r8), allocatable :: vp(:), vlat(:), vlon(:), hxa(:), el(:), ehl(:), ec(:), eclat(:), eclon(:)
    real(r8), allocatable :: etg(:), enr(:), ehc(:), ehnr(:), ehtg(:), tra(:), trc(:), trclat(:), trclon(:)
    integer :: i, l, k, n, o, nv, ne, nt

    call bin_read(fname)
    dims = geti('dims')
    nv = dims(1); ne = dims(2); nt = dims(3)
    mesh%nv = nv; mesh%ne = ne; mesh%nt = nt; mesh%glevel = dims(4)
    mesh%kind = 'icos'; mesh%pos = 'eqs'; mesh%optm = 'nopt'; mesh%hrchy = 0
    mesh%maxvnb = mx
    allocate(mesh%v(nv), mesh%hx(nv), mesh%ed(ne), mesh%edhx(ne), mesh%tr(nt))
    vp = getr('v_p'); vlat = getr('v_lat'); vlon = getr('v_lon')
    nnb = geti('v_nnb'); nb = geti('v_nb'); ved = geti('v_ed'); vtr = geti('v_tr')
    ttg = geti('hx_ttgout'); tnr = geti('hx_tnrccw'); htg = geti('hx_tg'); hnr = geti('hx_nr')
    hxa = getr('hx_areag')
    do i = 1, nv
      n = nnb(i); o = mx*(i-1)
      mesh%v(i)%p = vp(3*i-2:3*i)
      mesh%v(i)%lat = vlat(i); mesh%v(i)%lon = vlon(i)
      mesh%v(i)%nnb = n
      allocate(mesh%v(i)%nb(n), mesh%v(i)%ed(n), mesh%v(i)%tr(n))
      mesh%v(i)%nb = nb(o+1:o+n); mesh%v(i)%ed = ved(o+1:o+n); mesh%v(i)%tr = vtr(o+1:o+n)
      allocate(mesh%hx(i)%ttgout(n), mesh%hx(i)%tnrccw(n), mesh%hx(i)%tg(n), mesh%hx(i)%nr(n))
      mesh%hx(i)%ttgout = ttg(o+1:o+n); mesh%hx(i)%tnrccw = tnr(o+1:o+n)
      mesh%hx(i)%tg = htg(o+1:o+n); mesh%hx(i)%nr = hnr(o+1:o+n)
      mesh%hx(i)%areag = hxa(i)
    end do
    edv = geti('ed_v'); edsh = geti('ed_sh'); el = getr('ed_leng'); ehl = getr('edhx_leng')
    ec = getr('ed_c'); eclat = getr('ed_clat'); eclon = getr('ed_clon')
    etg = getr('ed_tg'); enr = getr('ed_nr'); ehc = getr('edhx_c'); ehnr = getr('edhx_nr'); ehtg = getr('edhx_tg')
    do l = 1, ne
      mesh%ed(l)%v = edv(2*l-1:2*l); mesh%ed(l)%sh = edsh(2*l-1:2*l)
      mesh%ed(l)%leng = el(l)
      mesh%ed(l)%c%p = ec(3*l-2:3*l); mesh%ed(l)%c%lat = eclat(l); mesh%ed(l)%c%lon = eclon(l)
      mesh%ed(l)%tg = etg(3*l-2:3*l); mesh%ed(l)%nr = enr(3*l-2:3*l)
      mesh%edhx(l)%sh = mesh%ed(l)%v; mesh%edhx(l)%v = mesh%ed(l)%sh
      mesh%edhx(l)%leng = ehl(l)
      mesh%edhx(l)%c%p = ehc(3*l-2:3*l)
      mesh%edhx(l)%nr = ehnr(3*l-2:3*l); mesh%edhx(l)%tg = ehtg(3*l-2:3*l)
    end do
    trv = geti('tr_v'); tred = geti('tr_ed'); trtg = geti('tr_tg'); trnr = geti('tr_nr')
    tra = getr('tr_areag'); trc = getr('tr_c'); trclat = getr('tr_clat'); trclon = getr('tr_clon')
    do k = 1, nt
      mesh%tr(k)%v = trv(3*k-2:3*k); mesh%tr(k)%ed = tred(3*k-2:3*k)
      mesh%tr(k)%tg = trtg(3*k-2:3*k); mesh%tr(k)%nr = trnr(3*k-2:3*k)
      mesh%tr(k)%areag = tra(k)
      mesh%tr(k)%c%p = trc(3*k-2:3*k); mesh%tr(k)%c%lat = trclat(k); mesh%tr(k)%c%lon = trclon(k)
    end do
  end subroutine load_mesh
end module harness_mesh

program ref_harness
  use constants, only: i4, r8, grav, erad
  use imxbin
  use harness_mesh
  use smeshpack, only: meshbuild, meshload, showonscreen
  use swm_data
  use swm, only: tendency, ode_rk4, initialize_gridprop, initialize_fields, &
       calc_energies, sumf_areas, sumfsq_areas, swm_horiz_loc_trunc_er, error_calc, error_calc_vec
  use omp_lib
  use interpack, only: plot_scalarfield
  use diffoperpack, only: grad_edge_Cgrid
  use datastruct, only: scalar_field, vector_field_cart, grid_structure
  use swm_operators, only: alignment_coef, diameter_coef, scalar_hx2ed, scalar_hx2tr
  implicit none

  character(len=512) :: cmd, a1, a2, a3, meshsrc, statesrc, outf
  type(scalar_field) :: pf
  character(len=128) :: hdr
  integer :: glevel, nsteps, idiag, k, ios, pc1, pc2
  real(r8) :: dtin, time, wall0, wall1
  real(r8) :: pe0, ke0, te0, ae0, pe, ke, te, ae, mass0
  real(r8), allocatable :: series(:, :)
  real(r8), allocatable :: hst(:), ust(:)
  integer :: pu, l
  type(grid_structure) :: meshcp
  character(len=512) :: scheme
  character(len=16) :: sc_stag = 'HTC', sc_wrap = 'trsk10', sc_diffus = 'const', sc_hdiffus = 'const'
  character(len=16) :: sc_pv = 'none'
  real(r8) :: sc_k2 = 0._r8, sc_k4 = 0._r8, sc_pvspar = 0.5_r8
  integer :: sc_tc = 2, sc_nopv = 0
  character(len=16) :: sc_rec = 'trsk', sc_cor = 'trsk', sc_sint = 'trsk'
  real(r8) :: sc_gcoef = 1.0_r8

  call get_command_argument(1, cmd)
  call execute_command_line('mkdir -p par grid data')
  showonscreen = .false.

  select case (trim(cmd))
  case ('build')
    call get_command_argument(2, a1); read(a1, *) glevel
    call get_command_argument(3, a2)
    call get_command_argument(4, outf)
    call build_ref_mesh(glevel, trim(a2))
    call write_swm_par(900._r8)
    call swmpars(usetime=.true.)
    call allocate_globalswmvars()
    call initialize_gridprop()
    call bin_open_w(trim(outf))
    call dump_mesh(mesh, .true.)
    call bin_close()

  case ('buildread')
    call get_command_argument(2, a1)
    call get_command_argument(3, a2); read(a2, *) glevel
    call get_command_argument(4, outf)
    mesh%kind = 'read'
    mesh%pos = 'eqs'
    mesh%optm = 'nopt'
    mesh%loadable = 0
    mesh%hrchy = 0
    mesh%nv = glevel
    mesh%name = trim(a1)
    call meshbuild(mesh)
    call write_swm_par(900._r8)
    call swmpars(usetime=.true.)
    call allocate_globalswmvars()
    call initialize_gridprop()
    call bin_open_w(trim(outf))
    call dump_mesh(mesh, .true.)
    call bin_close()

  case ('plot')
    ! ref_harness plot <glevel> <field.bin> <pos>: plot_scalarfield of the
    ! record 'f' (reference numbering) on the reference's glevel/eqs mesh,
    ! into data/plottest_<mesh name>.dat
    call get_command_argument(2, a1); read(a1, *) glevel
    call get_command_argument(3, a2)
    call get_command_argument(4, a3); read(a3, *) pc1
    call build_ref_mesh(glevel, 'eqs')
    call bin_read(trim(a2))
    pf%f = getr('f')
    pf%n = size(pf%f)
    pf%pos = pc1
    pf%name = 'plottest'
    call plot_scalarfield(pf, mesh)

  case ('loadref')
    ! ref_harness loadref <name> <out.bin>: meshload of grid/<name>_*.dat
    call get_command_argument(2, a1)
    call get_command_argument(3, outf)
    mesh%kind = 'icos'
    mesh%pos = 'eqs'   ! meshload leaves pos to the caller (printmesh checks it)
    mesh%optm = 'nopt'
    mesh%name = trim(a1)
    hdr = 'grid/'//trim(a1)//'_header.dat'
    call meshload(mesh, hdr)
    call bin_open_w(trim(outf))
    call dump_mesh(mesh, .false.)
    call bin_close()

  case ('run')
    call get_command_argument(2, meshsrc)
    call get_command_argument(3, statesrc)
    call get_command_argument(4, a1); read(a1, *) nsteps
    call get_command_argument(5, a2); read(a2, *) dtin
    call get_command_argument(6, a3); read(a3, *) idiag
    call get_command_argument(7, outf)
    call get_command_argument(8, scheme)
    call parse_scheme(trim(scheme))

    if (meshsrc(1:6) == 'build:') then
      pc1 = 7
      pc2 = index(meshsrc(pc1:), ':') + pc1 - 1
      read(meshsrc(pc1:pc2-1), *) glevel
      call build_ref_mesh(glevel, trim(meshsrc(pc2+1:)))
    else
      call load_mesh(trim(meshsrc), mesh)
    end if

    call write_swm_par(dtin)
    call swmpars(usetime=.true.)
    call allocate_globalswmvars()
    call initialize_gridprop()
    call initialize_fields()
    call setup_diffusion()
    if (trim(statesrc) /= 'tc2' .and. trim(statesrc) /= 'tc') then
      call bin_read(trim(statesrc))
      hst = getr('h'); ust = getr('u')
      if (size(hst) /= mesh%nv .or. size(ust) /= mesh%ne) stop 'run: state size mismatch'
      h%f = hst; u%f = ust
      bt%f = 0._r8
    end if

    call bin_open_w(trim(outf))
    call put_f8('h0', h%f); call put_f8('u0', u%f); call put_f8('bt', bt%f)

    !One tendency at the initial state (swm.f90:214) and its intermediates
    call tendency(h, u, masseq%f, momeq%f)
    call put_f8('masseq', masseq%f); call put_f8('momeq', momeq%f)
    call put_f8('h_ed', h_ed%f); call put_f8('uh', uh%f)
    call put_f8('eta', eta%f); call put_f8('h_tr', h_tr%f)
    call put_f8('q_tr', q_tr%f); call put_f8('q_ed', q_ed%f)
    call put_f8('divuh', divuh%f); call put_f8('ke_hx', ke_hx%f)
    call put_f8('ghbK', ghbK%f); call put_f8('grad_ghbK', grad_ghbK%f)
    call put_f8('uhq_perp', uhq_perp%f)
    call put_f8('divu', divu%f); call put_f8('zeta', zeta%f)
    call put_f8('v_hx', (/ (v_hx%p(k)%v, k = 1, mesh%nv) /))
    if (useAPVM) then
      call put_f8('q_grad_ed', q_grad_ed%f)
      call put_f8('q_grad_tr', (/ (q_grad_tr%p(k)%v, k = 1, mesh%nt) /))
    end if
    if (K2_max /= 0 .or. K4_max /= 0) then
      call put_f8('dif_coef_hx', dif_coef_hx%f); call put_f8('dif_coef_tr', dif_coef_tr%f)
      call put_f8('lapu', lapu%f)
    end if
    if (K4_max /= 0) call put_f8('lap_lapu', lap_lapu%f)

    !Initial diagnostics (swm.f90:224-227)
    mass0 = sumf_areas(h)
    call calc_energies(pe0, ke0, te0, ae0)
    call put_f8('diag0', (/ mass0, pe0, ke0, te0, ae0, dsqrt(sumfsq_areas(divu)) /))

    !Time loop of swm_tests (swm.f90:238-370) without output/reference branches
    allocate(series(7, max(nsteps, 1)))
    series = 0._r8
    u_old = u
    h_old = h
    wall0 = omp_get_wtime()
    do k = 1, nsteps
      time = real(k, r8)*dtin
      call ode_rk4(time, h_old, u_old, h, u, dtin)
      if (idiag == 1) then
        series(1, k) = time
        series(2, k) = sumf_areas(h)
        call calc_energies(pe, ke, te, ae)
        series(3, k) = pe; series(4, k) = ke; series(5, k) = te; series(6, k) = ae
        series(7, k) = dsqrt(sumfsq_areas(divu))
      end if
      u_old = u
      h_old = h
    end do
    wall1 = omp_get_wtime()
    call put_f8('h', h%f); call put_f8('u', u%f)
    if (idiag == 1) call put_f8('series', reshape(series, (/ 7*max(nsteps, 1) /)))
    call put_f8('wall', (/ wall1 - wall0, real(omp_get_max_threads(), r8) /))
    call bin_close()

  case ('lte')
    call get_command_argument(2, meshsrc)
    call get_command_argument(3, outf)
    call get_command_argument(4, scheme)
    call parse_scheme(trim(scheme))
    if (meshsrc(1:6) == 'build:') then
      pc1 = 7
      pc2 = index(meshsrc(pc1:), ':') + pc1 - 1
      read(meshsrc(pc1:pc2-1), *) glevel
      call build_ref_mesh(glevel, trim(meshsrc(pc2+1:)))
    else
      call load_mesh(trim(meshsrc), mesh)
    end if
    call write_swm_par(900._r8)
    !the routine assigns its argument to the module mesh: hand it a copy
    meshcp = mesh
    call swm_horiz_loc_trunc_er(meshcp)
    !it leaves grad_ghbK holding the gradient of the exact potential: the
    !same tendency again restores the fields of the rows it printed
    call tendency(h, u, masseq%f, momeq%f)
    call bin_open_w(trim(outf))
    call put_f8('h0', h%f); call put_f8('u0', u%f)
    call lte_s('lte_h_ed', h_ed, h_ed_exact, h_ed_error)
    call lte_s('lte_divuh', divuh, divuh_exact, divuh_error)
    call lte_s('lte_kenergy', ke_hx, ke_hx_exact, ke_hx_error)
    call lte_s('lte_abs_vort', eta, eta_exact, eta_error)
    call lte_s('lte_h_tr', h_tr, h_tr_exact, h_tr_error)
    call lte_s('lte_q_tr', q_tr, q_tr_exact, q_tr_error)
    call lte_s('lte_q_hx', q_hx, q_hx_exact, q_hx_error)
    call lte_s('lte_q_ed', q_ed, q_ed_exact, q_ed_error)
    call lte_s('lte_grad_ghbK', grad_ghbK, grad_ghbK_exact, grad_ghbK_error)
    call put_f8('x_grad_ghbK', grad_ghbK_exact%f)
    !the row of the gradient of the exact potential (src/swm.f90:548-558)
    ghbK%f(1:ghbK%n) = grav*(h_exact%f(1:ghbK%n) + bt%f(1:ghbK%n)) + ke_hx_exact%f(1:ghbK%n)
    do l = 1, mesh%ne
      grad_ghbK%f(l) = grad_edge_Cgrid(l, ghbk, u%pos, gradmtd, mesh)/erad
    end do
    call lte_s('lte_grad_ghbK_exact', grad_ghbK, grad_ghbK_exact, grad_ghbK_error)
    call lte_v('lte_vhq_tr', vhq_tr, vhq_tr_exact, vhq_tr_error)
    if (useAPVM) call lte_v('lte_q_grad_tr', q_grad_tr, q_grad_tr_exact, q_grad_tr_error)
    call lte_v('lte_vh_hx', vh_hx, vh_hx_exact, vh_hx_error)
    call lte_v('lte_vhq_hx', vhq_hx, vhq_hx_exact, vhq_hx_error)
    call lte_s('lte_uhq_perp', uhq_perp, uhq_perp_exact, uhq_perp_error)
    call lte_s('lte_masseq', masseq, masseq_exact, masseq_error)
    call lte_s('lte_momeq', momeq, momeq_exact, momeq_error)
    !fields of the tendency and the exact fields, for pinning each part
    call put_f8('vhq_tr', (/ (vhq_tr%p(k)%v, k = 1, mesh%nt) /))
    call put_f8('vh_hx', (/ (vh_hx%p(k)%v, k = 1, mesh%nv) /))
    call put_f8('q_hx', q_hx%f)
    call put_f8('x_h_ed', h_ed_exact%f); call put_f8('x_ke_hx', ke_hx_exact%f)
    call put_f8('x_eta', eta_exact%f); call put_f8('x_q_ed', q_ed_exact%f)
    call put_f8('x_uhq_perp', uhq_perp_exact%f); call put_f8('x_lapu', lapu_exact%f)
    call put_f8('x_q_hx', q_hx_exact%f)
    call put_f8('x_vhq_tr', (/ (vhq_tr_exact%p(k)%v, k = 1, mesh%nt) /))
    call put_f8('x_q_grad_tr', (/ (q_grad_tr_exact%p(k)%v, k = 1, mesh%nt) /))
    call put_f8('x_vhq_hx', (/ (vhq_hx_exact%p(k)%v, k = 1, mesh%nv) /))
    call bin_close()

  case default
    print*, 'usage: ref_harness build <glevel> <eqs|pol> <out.bin>'
    print*, '       ref_harness run <build:G:pos|mesh.bin> <tc|tc2|state.bin> <nsteps> <dt> <diag> <out.bin> [tc=N,...]'
    print*, '       ref_harness loadref <name> <out.bin>   (reads grid/<name>_*.dat)'
    print*, '       ref_harness plot <glevel> <field.bin> <pos>   (data/plottest_<grid>.dat)'
    stop 1
  end select

contains

  !error_calc of one scalar row of the LTE table (maxrel, rel L2, max)
  subroutine lte_s(name, f, fref, fdif)
    character(len=*), intent(in) :: name
    type(scalar_field), intent(in) :: f, fref
    type(scalar_field), intent(inout) :: fdif
    real(r8) :: a, b, c
    call error_calc(f, fref, fdif, a, b, c)
    call put_f8(name, (/ a, b, c /))
  end subroutine lte_s

  subroutine lte_v(name, v, vref, vdif)
    character(len=*), intent(in) :: name
    type(vector_field_cart), intent(in) :: v, vref
    type(vector_field_cart), intent(inout) :: vdif
    real(r8) :: a, b, c
    call error_calc_vec(v, vref, vdif, a, b, c)
    call put_f8(name, (/ a, b, c /))
  end subroutine lte_v

  subroutine build_ref_mesh(g, pos)
    integer, intent(in) :: g
    character(len=*), intent(in) :: pos
    mesh%kind = 'icos'
    mesh%pos = pos
    mesh%optm = 'nopt'
    mesh%loadable = 0
    mesh%hrchy = 0
    mesh%nv = 10*2**(2*g) + 2
    mesh%name = ''
    call meshbuild(mesh)
  end subroutine build_ref_mesh

  !key=value list of the scheme options written into par/swm.par
  subroutine parse_scheme(str)
    character(len=*), intent(in) :: str
    integer :: a, b, eq
    character(len=64) :: item, key, val
    a = 1
    do while (a <= len(str))
      b = index(str(a:), ',')
      if (b == 0) then
        b = len(str) + 1
      else
        b = a + b - 1
      end if
      item = str(a:b-1)
      a = b + 1
      eq = index(item, '=')
      if (eq == 0) cycle
      key = item(1:eq-1)
      val = item(eq+1:)
      select case (trim(key))
      case ('stag'); sc_stag = val
      case ('wrap'); sc_wrap = val
      case ('k2'); read(val, *) sc_k2
      case ('diffus'); sc_diffus = val
      case ('k4'); read(val, *) sc_k4
      case ('hyperdiffus'); sc_hdiffus = val
      case ('pv'); sc_pv = val
      case ('pvspar'); read(val, *) sc_pvspar
      case ('tc'); read(val, *) sc_tc
      case ('rec'); sc_rec = val
      case ('cor'); sc_cor = val
      case ('sint'); sc_sint = val
      case ('gasscoef'); read(val, *) sc_gcoef
      case ('nopv'); read(val, *) sc_nopv
      case default
        print*, 'unknown scheme key ', trim(key)
        stop 1
      end select
    end do
  end subroutine parse_scheme

  !Diffusion and hyperdiffusion coefficients as swm_tests sets them up
  !after initialize_fields (src/swm.f90:170-210)
  subroutine setup_diffusion()
    if (K2_max > 0.d0) then
      select case (diffus)
      case ('const'); dif_coef_hx%f = K2_max
      case ('align'); call alignment_coef(K2_max, dif_coef_hx, mesh)
      case ('diam'); call diameter_coef(K2_max, dif_coef_hx, mesh)
      case default; stop 'invalid diffusion coefficient function'
      end select
      call scalar_hx2ed(dif_coef_hx, dif_coef_ed, mesh)
      call scalar_hx2tr(dif_coef_hx, dif_coef_tr, mesh)
    end if
    if (K4_max > 0.d0) then
      select case (hyperdiffus)
      case ('const'); hypdif_coef_hx%f = K4_max
      case ('align'); call alignment_coef(K4_max, hypdif_coef_hx, mesh)
      case ('diam'); call diameter_coef(K4_max, hypdif_coef_hx, mesh)
      case default; stop 'invalid hyperdiffusion coefficient function'
      end select
      call scalar_hx2ed(hypdif_coef_hx, hypdif_coef_ed, mesh)
      call scalar_hx2tr(hypdif_coef_hx, hypdif_coef_tr, mesh)
      dif_coef_hx%f = dsqrt(hypdif_coef_hx%f)
      dif_coef_ed%f = dsqrt(hypdif_coef_ed%f)
      dif_coef_tr%f = dsqrt(hypdif_coef_tr%f)
    end if
  end subroutine setup_diffusion

  !Positional par/swm.par in the layout read by swmpars (swm_data.f90:389-424):
  !TC2 (overridden by a state file when given), the scheme options of
  !parse_scheme, no plots.
  subroutine write_swm_par(dtv)
    real(r8), intent(in) :: dtv
    open(newunit=pu, file='par/swm.par', status='replace', iostat=ios)
    write(pu, '(a)') '!c'
    write(pu, '(a)') '!c'
    write(pu, '(i0,a)') sc_tc, ' 0'
    write(pu, '(a)') '!c'
    write(pu, '(a)') '1 1'
    write(pu, '(a)') '!c'
    write(pu, '(es25.17,a)') dtv, ' 0 0'
    write(pu, '(a)') '!c'
    write(pu, '(a)') trim(sc_stag)
    write(pu, '(a)') '!c'
    write(pu, '(a)') trim(sc_wrap)
    write(pu, '(a)') '!c'
    write(pu, '(a,1x,es25.17)') trim(sc_rec), sc_gcoef
    write(pu, '(a)') '!c'
    write(pu, '(a)') trim(sc_cor)
    write(pu, '(a)') '!c'
    write(pu, '(a)') trim(sc_sint)
    write(pu, '(a)') '!c'
    write(pu, '(a)') 'trsk'
    write(pu, '(a)') '!c'
    write(pu, '(a)') 'geo'
    write(pu, '(a)') '!c'
    write(pu, '(a)') '0 1000000 0'
    write(pu, '(a)') '!c'
    write(pu, '(es25.17,1x,a)') sc_k2, trim(sc_diffus)
    write(pu, '(a)') '!c'
    write(pu, '(es25.17,1x,a)') sc_k4, trim(sc_hdiffus)
    write(pu, '(a)') '!c'
    write(pu, '(a)') '2.0'
    write(pu, '(a)') '!c'
    write(pu, '(a)') '0.0'
    write(pu, '(a)') '!c'
    write(pu, '(a,1x,es25.17)') trim(sc_pv), sc_pvspar
    write(pu, '(a)') '!c'
    write(pu, '(i0)') sc_nopv
    close(pu)
  end subroutine write_swm_par
end program ref_harness
