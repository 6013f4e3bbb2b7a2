! lepic_driver.f90 -- a Fortran time loop over the GPU hot path.
!
! Reads a namelist (&lepic ... /, defaults below give a grounded box with
! e-/Ar+ at ppc particles per cell, a Gaussian Bz filter and analytic Ar
! cross sections), then runs the reference's per-step order (main.f90:689-
! 1090): rhs -> MG Poisson -> E -> sort every nsort steps -> density
! reduction + MCC every ns_coll steps -> push + deposition of each species,
! each phase through the C ABI; averaged moments every navg steps on the grid
! coarsened by size_na (main.f90:903-944), particle backups every nbak steps
! into DATA.BAK/particles0.bak (main.f90:1102-1132; the directory must
! exist, as for the reference), and restart = 1 starts
! from that file (restart.f90:56-64) instead of a uniform load.  With
! flag_heat = 1 the electrons in the slit or disk of e_heating.f90 are heated
! every ns_heat steps before the push (main.f90:1033-1037); xa and dr default
! to the centre and a quarter of the box, the slit to the box height.  Prints the
! phase timers in the style of the reference's "<t> (ms)" lines
! (main.f90:1318-1322), the particle counts and the averaged Te at the centre.
program lepic_driver
  use iso_c_binding
  use lepic_capi
  implicit none
  ! ---- namelist ----
  integer :: n = 128, ppc = 20, nsteps = 20, ng = 7, precision = 0, nsort = 10, ns_coll = 10
  integer :: mcc_mode = 1, navg = 0, size_na = 1, nbak = 0, restart = 0
  integer :: flag_heat = 0, ns_heat = 1, flag_c = 1, flag_gpp = 0, eheat_type = 0
  real(8) :: xa = -1.d0, dr = -1.d0, yl_pow = 0.d0, yr_pow = -1.d0, nudt = 1.d-3
  real(8) :: dx = 5.d-5, dt = 5.d-12, Te = 3.d0, Ti = 0.03d0, Tgas = 0.02585d0, n0 = 1.d16
  real(8) :: B0 = 0.01d0, pgas = 0.3d0, Tgas_K = 300.d0, eps = 1.d-10
  namelist /lepic/ n, ppc, nsteps, ng, precision, nsort, ns_coll, mcc_mode, dx, dt, Te, Ti, Tgas, &
       n0, B0, pgas, Tgas_K, eps, navg, size_na, nbak, restart, flag_heat, ns_heat, flag_c, flag_gpp, &
       eheat_type, xa, dr, yl_pow, yr_pow, nudt
  ! ---- constants (main.f90:120-124) ----
  real(8) :: qe, amu, eps0, pi, me, mAr, kB
  ! ---- state ----
  type(c_ptr) :: ctx
  type(lepic_grid_desc) :: g
  type(lepic_species_desc) :: sp(2)
  type(lepic_solve_info) :: info
  type(lepic_mcc_info) :: minfo
  type(lepic_reactions) :: rx
  type(lepic_heating) :: heat
  real(8), target :: vt_heat(1)
  real(8) :: dP, P_heat
  integer(c_int32_t), allocatable, target :: bcnd(:,:)
  real(8), allocatable, target :: phi(:,:), Bi(:,:,:)
  real(8), allocatable, target :: Er(:), sig(:,:), Eth(:), sigv(:,:), charge(:), mass(:), Nm(:)
  real(8), allocatable, target :: vt0(:), np_mx(:), nu_uplim(:)
  integer(c_int32_t), allocatable, target :: col_info(:,:), p_ncol(:), sig_list(:,:)
  real(8) :: ngas, Nmw, L, x, vte, vti, vtg, mu, vr, np_mx_c(8)
  real(8) :: t_field, t_sort, t_mcc, t_move
  integer(c_int64_t) :: ne, ni
  integer :: it, i, ix, u, ios, npt, c, rc, vcyc
  integer(c_int32_t) :: iseed
  integer(8) :: c0, c1, crate
  real(8), allocatable :: pavg(:,:,:,:), phiavg(:,:,:)
  real(8) :: time, pabs_cor
  integer(c_int32_t) :: gtype, cnt_avg
  character(len=*), parameter :: bakfile = 'DATA.BAK/particles0.bak'

  open(newunit=u, file='lepic.nml', status='old', iostat=ios)
  if (ios == 0) then
     read(u, nml=lepic)
     close(u)
  end if
  qe = real(1.60217646e-19, 8)
  amu = 1.66053886d-27
  eps0 = 8.854187817d-12
  pi = 4.d0*datan(1.d0)
  me = 9.10938188d-31
  mAr = 39.95d0*amu
  kB = 1.380649d-23
  ngas = pgas/(kB*Tgas_K)
  Nmw = n0*dx*dx*1.d0/ppc
  L = n*dx
  vte = dsqrt(2.d0*qe*Te/me)
  vti = dsqrt(2.d0*qe*Ti/mAr)
  vtg = dsqrt(2.d0*qe*Tgas/mAr)

  g = lepic_grid_desc(n, n, dx, dx, 1.d0, dt, eps0, qe, ng, 1, 2, precision, 0, 0, 0, 0)
  sp(1) = lepic_species_desc(-qe, me, Nmw, int(n, 8)*n*ppc*2 + 4096)
  sp(2) = lepic_species_desc(qe, mAr, Nmw, int(n, 8)*n*ppc*2 + 4096)
  call lepic_check(ctx, lepic_create(g, sp, ctx), 'create')

  ! grounded box (generate_boundary.f90 drawn on the four sides)
  allocate(bcnd(0:n+2, 0:n+2), phi(0:n+2, 0:n+2), Bi(4, 0:n+2, 0:n+2))
  bcnd = -1
  bcnd(0:1, :) = 1; bcnd(n+1:n+2, :) = 1; bcnd(:, 0:1) = 1; bcnd(:, n+1:n+2) = 1
  phi = 0.d0
  call lepic_check(ctx, lepic_set_boundary(ctx, bcnd, phi), 'set_boundary')
  ! Gaussian Bz(x) filter (Bfield.f90:53-59) and |B| (main.f90:411-422)
  Bi = 0.d0
  do ix = 0, n + 2
     x = (ix - 1)*dx
     Bi(3, ix, :) = B0*dexp(-(x - L/2)**2/(2*(L/8)**2))
  end do
  Bi(4, :, :) = dabs(Bi(3, :, :))
  call lepic_check(ctx, lepic_set_bfield(ctx, n, n, dx, dx, Bi), 'set_bfield')
  time = 0.d0; gtype = 0; pabs_cor = 1.d0
  if (restart == 1) then
     call lepic_check(ctx, lepic_load_particles(ctx, bakfile//c_null_char, time, gtype, pabs_cor), 'restart')
  else
     call lepic_check(ctx, lepic_load_uniform(ctx, 1, ppc, vte, 0.d0, L, 20261202_8), 'load e')
     call lepic_check(ctx, lepic_load_uniform(ctx, 2, ppc, vti, 0.d0, L, 20261202_8), 'load Ar+')
  end if

  ! analytic Ar table on one energy grid (union grid = the grid itself)
  npt = 242
  allocate(Er(npt), Eth(3), sigv(3, 0:3), charge(0:3), mass(0:3), Nm(0:3))
  allocate(vt0(0:3), np_mx(0:3), nu_uplim(0:3), col_info(10, 3), p_ncol(0:3), sig_list(3, 0:3))
  Er(1) = 0.d0
  do i = 2, npt - 2
     Er(i) = 1.d-3*10.d0**(7.d0*(i - 2)/(npt - 4))
  end do
  Er(npt - 1) = 11.5d0
  Er(npt) = 15.76d0
  call sort_unique(Er, npt)
  allocate(sig(npt, 3))
  do i = 1, npt
     sig(i, 1) = 1.d-19
     sig(i, 2) = 0.d0
     if (Er(i) > 11.5d0) sig(i, 2) = 1.d-20*(1.d0 - 11.5d0/Er(i))*dlog(Er(i)/11.5d0)/(Er(i)/11.5d0)*dexp(1.d0)
     sig(i, 3) = 0.d0
     if (Er(i) > 15.76d0) sig(i, 3) = 4.d-18*6.d0*dlog(Er(i)/15.76d0)/(Er(i)*15.76d0)* &
          (1.d0 - 0.62d0*dexp(-0.4d0*(Er(i)/15.76d0 - 1.d0)))
  end do
  Eth = [0.d0, 11.5d0, 15.76d0]
  col_info = 0
  col_info(1:7, 1) = [2, 2, 1, 3, 1, 3, 1]          ! [e]+[Ar]->[e]+[Ar] COLLISION
  col_info(1:7, 2) = [2, 2, 1, 3, 1, 3, 3]          ! EXCITATION
  col_info(1:8, 3) = [2, 3, 1, 3, 1, 2, 1, 2]       ! [e]+[Ar]->[e]+[Ar+]+[e] IONIZATION
  col_info(9, 3) = 2
  charge = [0.d0, -qe, qe, 0.d0]
  mass = [0.d0, me, mAr, -mAr]
  Nm = [0.d0, Nmw, Nmw, ngas*dx*dx/ppc]
  vt0 = [0.d0, vte, vti, vtg]
  np_mx = [0.d0, 0.d0, 0.d0, ngas]
  nu_uplim = [0.d0, 6.d7, 5.d6, 5.d6]
  p_ncol = [0, 3, 0, 0]
  sig_list = 0
  sig_list(:, 1) = [1, 2, 3]
  sigv = 0.d0                                      ! read_reactions.f90:657-672
  mu = me*mAr/(me + mAr)
  do c = 1, 3
     do i = 1, npt
        vr = dsqrt(2.d0*Er(i)*qe/mu)
        if (Er(i) >= 0.1d0) sigv(c, 1) = max(sigv(c, 1), vr*sig(i, c))
     end do
  end do
  rx%ntype_all = 3; rx%ncol = 3; rx%npt = npt; rx%ns_coll = ns_coll; rx%n_cell = n*n
  rx%n_rg = 1; rx%iyl_rg = 1; rx%iyr_rg = n + 1; rx%ixl_rg = 0; rx%ixr_rg = 0
  rx%ixl_rg(1) = 1; rx%ixr_rg(1) = n + 1; rx%flag_ionrep = 0; rx%reserved = 0
  rx%sig_Er = c_loc(Er(1)); rx%sig = c_loc(sig); rx%Eth = c_loc(Eth); rx%sigv_mx = c_loc(sigv)
  rx%col_info = c_loc(col_info); rx%p_ncol = c_loc(p_ncol); rx%sig_list = c_loc(sig_list)
  rx%charge = c_loc(charge); rx%mass = c_loc(mass); rx%Nm = c_loc(Nm); rx%vt0 = c_loc(vt0)
  rx%np_mx = c_loc(np_mx); rx%nu_uplim = c_loc(nu_uplim)
  call lepic_check(ctx, lepic_set_reactions(ctx, rx), 'set_reactions')
  do i = 1, 2
     call lepic_check(ctx, lepic_charge_deposition(ctx, i), 'deposit')
  end do

  ! electron heating parameters (particle_info.h members read by e_heating.f90)
  if (xa < 0.d0) xa = L/2
  if (dr < 0.d0) dr = L/4
  if (yr_pow < 0.d0) yr_pow = L
  vt_heat(1) = vte
  heat = lepic_heating(flag_c, flag_gpp, eheat_type, 1, xa, dr, yl_pow, yr_pow, nudt, pabs_cor, &
       c_loc(vt_heat(1)))
  P_heat = 0.d0

  ! ---- time loop (main.f90:689-1090) ----
  t_field = 0; t_sort = 0; t_mcc = 0; t_move = 0; vcyc = 0
  iseed = 123456
  call system_clock(count_rate=crate)
  do it = 1, nsteps
     call system_clock(c0)
     call lepic_check(ctx, lepic_calc_rhs(ctx), 'calc_rhs')
     if (it == 1) then
        rc = lepic_iterative_solver(ctx, eps, 20000, LEPIC_START_COLD, info, c_null_ptr, 0)
     else
        rc = lepic_iterative_solver(ctx, eps, 20000, LEPIC_START_PREV, info, c_null_ptr, 0)
     end if
     call lepic_check(ctx, rc, 'iterative_solver')
     vcyc = vcyc + info%vcycles
     call lepic_check(ctx, lepic_calc_efield(ctx), 'calc_Efield')
     call lepic_check(ctx, lepic_sync(ctx), 'sync')
     call system_clock(c1); t_field = t_field + real(c1 - c0, 8)/crate; c0 = c1
     if (mod(it, nsort) == 1) then
        do i = 1, 2
           call lepic_check(ctx, lepic_part_sorting(ctx, i, c_null_ptr), 'part_sorting')
        end do
        call lepic_check(ctx, lepic_sync(ctx), 'sync')
     end if
     call system_clock(c1); t_sort = t_sort + real(c1 - c0, 8)/crate; c0 = c1
     if (navg > 0) then
        if (mod(it, navg) == 1) call lepic_check(ctx, lepic_moments_accumulate(ctx, size_na), 'moments')
     end if
     if (it > 1 .and. mod(it, ns_coll) == 1) then
        call lepic_check(ctx, lepic_dens_red(ctx, np_mx_c), 'dens_red')
        call lepic_check(ctx, lepic_collisions(ctx, it, mcc_mode, 7_8, iseed, minfo), 'collisions')
     end if
     call system_clock(c1); t_mcc = t_mcc + real(c1 - c0, 8)/crate; c0 = c1
     if (flag_heat == 1 .and. mod(it, ns_heat) == 0) then
        call lepic_check(ctx, lepic_eheating(ctx, heat, mcc_mode, 11_8, it, iseed, dP), 'eheating')
        P_heat = P_heat + dP
     end if
     do i = 1, 2
        call lepic_check(ctx, lepic_part_mover(ctx, i, 1), 'part_mover')
     end do
     call lepic_check(ctx, lepic_sync(ctx), 'sync')
     call system_clock(c1); t_move = t_move + real(c1 - c0, 8)/crate
     time = time + dt
     if (nbak > 0 .and. it > 1) then
        if (mod(it, nbak) == 1) then
           call lepic_check(ctx, lepic_save_particles(ctx, bakfile//c_null_char, time, gtype, pabs_cor), 'backup')
           print '(a,i8)', ' backup written at step ', it
        end if
     end if
  end do
  call lepic_check(ctx, lepic_get_count(ctx, 1, ne), 'count')
  call lepic_check(ctx, lepic_get_count(ctx, 2, ni), 'count')
  print '(a,i12,a,i12)', ' particles: e- ', ne, '  Ar+ ', ni
  print '(a,f10.3)', ' V-cycles per step: ', real(vcyc, 8)/nsteps
  if (flag_heat == 1) print '(a,es14.6)', ' heating P_loss(2,1) (J): ', P_heat
  print '(a)', ' <t> (ms): E/rho+Poisson   sorting   MCC   mover'
  print '(a,4f12.4)', ' <t> (ms): ', 1.d3*t_field/nsteps, 1.d3*t_sort/nsteps, 1.d3*t_mcc/nsteps, &
       1.d3*t_move/nsteps
  call lepic_check(ctx, lepic_get_grid(ctx, LEPIC_GRID_PHI, 0, phi), 'get_grid')
  print '(a,es14.6)', ' phi(centre) (V): ', phi(n/2 + 1, n/2 + 1)
  if (navg > 0) then
     allocate(pavg(7, n/size_na + 1, n/size_na + 1, 2), phiavg(3, 0:n+2, 0:n+2))
     call lepic_check(ctx, lepic_moments_get(ctx, pavg, phiavg, cnt_avg), 'moments_get')
     if (cnt_avg > 0) print '(a,i6,a,es14.6)', ' averages: ', cnt_avg, '  Te(centre) (eV): ', &
          pavg(2, n/(2*size_na) + 1, n/(2*size_na) + 1, 1)/cnt_avg
  end if
  rc = lepic_destroy(ctx)

contains

  subroutine sort_unique(a, m)
    real(8), intent(inout) :: a(:)
    integer, intent(inout) :: m
    real(8) :: t
    integer :: p, q, k
    do p = 2, m
       t = a(p)
       q = p - 1
       do while (q >= 1)
          if (a(q) <= t) exit
          a(q + 1) = a(q)
          q = q - 1
       end do
       a(q + 1) = t
    end do
    k = 1
    do p = 2, m
       if (a(p) /= a(k)) then
          k = k + 1
          a(k) = a(p)
       end if
    end do
    m = k
  end subroutine sort_unique

end program lepic_driver
