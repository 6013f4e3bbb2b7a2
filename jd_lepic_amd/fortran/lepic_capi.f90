! lepic_capi.f90 -- iso_c_binding interfaces to the C ABI of liblepic_hip.so
! (include/lepic.h).  A Fortran time loop uses these in place of the
! reference's implicit-interface calls (main.f90:821-1086).
module lepic_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: LEPIC_FP64 = 0, LEPIC_MIXED = 1
  integer(c_int32_t), parameter :: LEPIC_MCC_CHECK = 0, LEPIC_MCC_PROD = 1
  integer(c_int32_t), parameter :: LEPIC_START_PREV = 0, LEPIC_START_COLD = 1, LEPIC_START_GIVEN = 2, &
       LEPIC_START_NET = 3
  integer(c_int32_t), parameter :: LEPIC_GRID_PHI = 0, LEPIC_GRID_RHS = 1, LEPIC_GRID_E = 2, &
       LEPIC_GRID_NP = 3, LEPIC_GRID_NPRED = 4, LEPIC_GRID_SINK = 5, LEPIC_GRID_SRC = 6

  type, bind(C) :: lepic_grid_desc
     integer(c_int32_t) :: n1, n2
     real(c_double) :: h1, h2, zmax, dt, eps0, qe
     integer(c_int32_t) :: ng, ngrid, nspecies, precision, flag_B_pos, tag_neg, track_z, mg_tail_nodes
  end type lepic_grid_desc

  type, bind(C) :: lepic_species_desc
     real(c_double) :: charge, mass, Nm
     integer(c_int64_t) :: capacity
  end type lepic_species_desc

  type, bind(C) :: lepic_solve_info
     integer(c_int32_t) :: vcycles, status
     real(c_double) :: res, ktot, res0
     integer(c_int32_t) :: start_used, nlog
     real(c_double) :: res0_alt
  end type lepic_solve_info

  type, bind(C) :: lepic_mcc_info
     integer(c_int64_t) :: Nc(8)
     real(c_double) :: nu_max(8)
     integer(c_int64_t) :: events(64), created(8), missed
     integer(c_int32_t) :: seed, reserved
  end type lepic_mcc_info

  type, bind(C) :: lepic_reactions
     integer(c_int32_t) :: ntype_all, ncol, npt, ns_coll, n_cell, n_rg, iyl_rg, iyr_rg
     integer(c_int32_t) :: ixl_rg(5), ixr_rg(5)
     integer(c_int32_t) :: flag_ionrep, reserved
     type(c_ptr) :: sig_Er, sig, Eth, sigv_mx
     type(c_ptr) :: col_info, p_ncol, sig_list
     type(c_ptr) :: charge, mass, Nm, vt0, np_mx, nu_uplim
  end type lepic_reactions

  type, bind(C) :: lepic_heating
     integer(c_int32_t) :: flag_c, flag_gpp, eheat_type, n_icp
     real(c_double) :: xa, dr, yl_pow, yr_pow, nudt, Pabs_cor
     type(c_ptr) :: vt
  end type lepic_heating

  type, bind(C) :: lepic_emitter
     real(c_double) :: jne, Lg, Lh, xg1, THm, chmfrd_array(4)
     integer(c_int32_t) :: n_holes, n_icp, ptype_fx, tag_b, flag_chmfrd, ns_Hm
  end type lepic_emitter

  type, bind(C) :: lepic_injection
     integer(c_int32_t) :: opt_inj, flag_c, flag_nmn, flag_Tp, flag_diffsrc, ns_inj
     real(c_double) :: I_inj, xl_pow, xr_pow, yl_pow, yr_pow, xa, dr, THm, th_B
     type(c_ptr) :: ni0, vt0
  end type lepic_injection

  type, bind(C) :: lepic_injection_info
     integer(c_int64_t) :: injected(8)
     real(c_double) :: P_loss4(8), P_loss1(8)
     integer(c_int32_t) :: seed, reserved
  end type lepic_injection_info

  type, bind(C) :: lepic_emitter_info
     integer(c_int64_t) :: Nh, created(8), cnt_b(2)
     real(c_double) :: P_loss4(8), Lgrd, dNh
     integer(c_int32_t) :: seed, reserved
  end type lepic_emitter_info

  interface
     integer(c_int) function lepic_create(g, sp, ctx) bind(C, name='lepic_create')
       import :: c_int, c_ptr, lepic_grid_desc, lepic_species_desc
       type(lepic_grid_desc), intent(in) :: g
       type(lepic_species_desc), intent(in) :: sp(*)
       type(c_ptr), intent(out) :: ctx
     end function lepic_create
     integer(c_int) function lepic_destroy(ctx) bind(C, name='lepic_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function lepic_destroy
     type(c_ptr) function lepic_last_error(ctx) bind(C, name='lepic_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function lepic_last_error
     integer(c_int) function lepic_sync(ctx) bind(C, name='lepic_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function lepic_sync
     integer(c_int) function lepic_set_boundary(ctx, bcnd, phi) bind(C, name='lepic_set_boundary')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: bcnd(*)
       real(c_double), intent(in) :: phi(*)
     end function lepic_set_boundary
     integer(c_int) function lepic_set_bfield(ctx, nB1, nB2, hB1, hB2, Bi) bind(C, name='lepic_set_bfield')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nB1, nB2
       real(c_double), value :: hB1, hB2
       real(c_double), intent(in) :: Bi(*)
     end function lepic_set_bfield
     integer(c_int) function lepic_load_uniform(ctx, ptype, ppc, vt, xl, xr, seed) &
          bind(C, name='lepic_load_uniform')
       import :: c_int, c_ptr, c_int32_t, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype, ppc
       real(c_double), value :: vt, xl, xr
       integer(c_int64_t), value :: seed
     end function lepic_load_uniform
     integer(c_int) function lepic_upload_particles(ctx, ptype, vxp, np) &
          bind(C, name='lepic_upload_particles')
       import :: c_int, c_ptr, c_int32_t, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype
       real(c_double), intent(in) :: vxp(*)
       integer(c_int64_t), value :: np
     end function lepic_upload_particles
     integer(c_int) function lepic_download_particles(ctx, ptype, vxp, cap, np) &
          bind(C, name='lepic_download_particles')
       import :: c_int, c_ptr, c_int32_t, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype
       real(c_double), intent(out) :: vxp(*)
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: np
     end function lepic_download_particles
     integer(c_int) function lepic_get_mover_tally(ctx, ptype, P_loss, p_mac, cnt_dead) &
          bind(C, name='lepic_get_mover_tally')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype
       real(c_double), intent(out) :: P_loss(*), p_mac(*), cnt_dead
     end function lepic_get_mover_tally
     integer(c_int) function lepic_get_count(ctx, ptype, np) bind(C, name='lepic_get_count')
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype
       integer(c_int64_t), intent(out) :: np
     end function lepic_get_count
     integer(c_int) function lepic_get_grid(ctx, which, ptype, out) bind(C, name='lepic_get_grid')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which, ptype
       real(c_double), intent(out) :: out(*)
     end function lepic_get_grid
     integer(c_int) function lepic_part_mover(ctx, ptype, deposit) bind(C, name='lepic_part_mover')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype, deposit
     end function lepic_part_mover
     integer(c_int) function lepic_charge_deposition(ctx, ptype) bind(C, name='lepic_charge_deposition')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype
     end function lepic_charge_deposition
     integer(c_int) function lepic_calc_rhs(ctx) bind(C, name='lepic_calc_rhs')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function lepic_calc_rhs
     integer(c_int) function lepic_calc_efield(ctx) bind(C, name='lepic_calc_efield')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function lepic_calc_efield
     integer(c_int) function lepic_iterative_solver(ctx, eps, ncycl, start, info, log, log_cap) &
          bind(C, name='lepic_iterative_solver')
       import :: c_int, c_ptr, c_int32_t, c_double, lepic_solve_info
       type(c_ptr), value :: ctx
       real(c_double), value :: eps
       integer(c_int32_t), value :: ncycl, start
       type(lepic_solve_info), intent(out) :: info
       type(c_ptr), value :: log
       integer(c_int32_t), value :: log_cap
     end function lepic_iterative_solver
     integer(c_int) function lepic_part_sorting(ctx, ptype, plist) bind(C, name='lepic_part_sorting')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ptype
       type(c_ptr), value :: plist
     end function lepic_part_sorting
     integer(c_int) function lepic_dens_red(ctx, np_mx) bind(C, name='lepic_dens_red')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: np_mx(*)
     end function lepic_dens_red
     integer(c_int) function lepic_set_reactions(ctx, r) bind(C, name='lepic_set_reactions')
       import :: c_int, c_ptr, lepic_reactions
       type(c_ptr), value :: ctx
       type(lepic_reactions), intent(in) :: r
     end function lepic_set_reactions
     integer(c_int) function lepic_collisions(ctx, it, mode, seed, iseed, info) bind(C, name='lepic_collisions')
       import :: c_int, c_ptr, c_int32_t, c_int64_t, lepic_mcc_info
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it, mode
       integer(c_int64_t), value :: seed
       integer(c_int32_t), intent(inout) :: iseed
       type(lepic_mcc_info), intent(out) :: info
     end function lepic_collisions
     ! electron heating (e_heating.f90:1-112), called where main.f90:1033-1037 calls eheating
     integer(c_int) function lepic_eheating(ctx, h, mode, seed, it, iseed, dP) bind(C, name='lepic_eheating')
       import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, lepic_heating
       type(c_ptr), value :: ctx
       type(lepic_heating), intent(in) :: h
       integer(c_int32_t), value :: mode, it
       integer(c_int64_t), value :: seed
       integer(c_int32_t), intent(inout) :: iseed
       real(c_double), intent(out) :: dP
     end function lepic_eheating
     ! surface emitter (part_flux_emitter.f90:1-217), called where main.f90:1092-1096 calls it
     integer(c_int) function lepic_flux_emitter(ctx, e, it, mode, seed, iseed, info) &
          bind(C, name='lepic_flux_emitter')
       import :: c_int, c_ptr, c_int32_t, c_int64_t, lepic_emitter, lepic_emitter_info
       type(c_ptr), value :: ctx
       type(lepic_emitter), intent(in) :: e
       integer(c_int32_t), value :: it, mode
       integer(c_int64_t), value :: seed
       integer(c_int32_t), intent(inout) :: iseed
       type(lepic_emitter_info), intent(out) :: info
     end function lepic_flux_emitter
     ! volume injection (part_injection.f90:1-237), the species loop of main.f90:1063-1071
     integer(c_int) function lepic_part_injection(ctx, inj, N_inj, it, mode, seed, iseed, deposit, info) &
          bind(C, name='lepic_part_injection')
       import :: c_int, c_ptr, c_int32_t, c_int64_t, lepic_injection, lepic_injection_info
       type(c_ptr), value :: ctx
       type(lepic_injection), intent(in) :: inj
       integer(c_int64_t), intent(in) :: N_inj(*)
       integer(c_int32_t), value :: it, mode, deposit
       integer(c_int64_t), value :: seed
       integer(c_int32_t), intent(inout) :: iseed
       type(lepic_injection_info), intent(out) :: info
     end function lepic_part_injection
     ! averaged moments (utils.f90:1-265) and backup files (main.f90:1102-1132, restart.f90)
     integer(c_int) function lepic_moments_accumulate(ctx, size_na) bind(C, name='lepic_moments_accumulate')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: size_na
     end function lepic_moments_accumulate
     integer(c_int) function lepic_moments_get(ctx, data_pavg, phi_avg, cnt) bind(C, name='lepic_moments_get')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: data_pavg(*), phi_avg(*)
       integer(c_int32_t), intent(out) :: cnt
     end function lepic_moments_get
     integer(c_int) function lepic_save_particles(ctx, path, time, gtype, pabs_cor) &
          bind(C, name='lepic_save_particles')
       import :: c_int, c_ptr, c_char, c_int32_t, c_double
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), value :: time, pabs_cor
       integer(c_int32_t), value :: gtype
     end function lepic_save_particles
     integer(c_int) function lepic_load_particles(ctx, path, time, gtype, pabs_cor) &
          bind(C, name='lepic_load_particles')
       import :: c_int, c_ptr, c_char, c_int32_t, c_double
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), intent(out) :: time, pabs_cor
       integer(c_int32_t), intent(out) :: gtype
     end function lepic_load_particles
     ! learned warm start (no counterpart in the reference)
     integer(c_int) function lepic_net_create(ctx, capacity, seed) bind(C, name='lepic_net_create')
       import :: c_int, c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: capacity
       integer(c_int64_t), value :: seed
     end function lepic_net_create
     integer(c_int) function lepic_net_recording(ctx, on) bind(C, name='lepic_net_recording')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on
     end function lepic_net_recording
     integer(c_int) function lepic_net_policy(ctx, min_gain, max_backoff) bind(C, name='lepic_net_policy')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       real(c_float), value :: min_gain
       integer(c_int32_t), value :: max_backoff
     end function lepic_net_policy
     integer(c_int) function lepic_net_train(ctx, iters, batch, lr, loss_first, loss_last) &
          bind(C, name='lepic_net_train')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: iters, batch
       real(c_float), value :: lr
       real(c_float), intent(out) :: loss_first, loss_last
     end function lepic_net_train
  end interface

contains

  ! stop with the library's message, like stop_calculation (restart.f90:263-279)
  subroutine lepic_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(lepic_last_error(ctx), msg, [512])
    n = 0
    do while (n < 512)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    print '(a,a,i4,a,512a)', where, ': lepic error ', rc, ': ', msg(1:n)
    stop 1
  end subroutine lepic_check

end module lepic_capi
