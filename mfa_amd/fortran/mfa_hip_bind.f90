! iso_c_binding interface to libmfa_hip (include/mfa_hip.h).
!
! The derived types mirror the C structs field by field (bind(C) types follow
! the C layout rules); mfa_check_abi() compares their sizes with the
! library's at start-up.  C `double x[4][4]` (row-major, [i][j]) appears here
! as x(4,4) with x(j,i) == C x[i][j].
module mfa_hip_bind
  use iso_c_binding
  implicit none

  integer, parameter :: MFA_NT = 4
  integer(c_int32_t), parameter :: MFA_FP64 = 0, MFA_MIXED = 1

  type, bind(C) :: mfa_inputs_t
    integer(c_int32_t) :: n_mon = 0, n_chain = 0, n_cell_w_x = 0, n_cell_w_y = 0
    integer(c_int32_t) :: n_mon_d = 0, n_chain_d = 0
    integer(c_int32_t) :: n_mon_e = 0, n_chain_e = 0, inter4 = 0, pad0 = 0
    real(c_double) :: a_w = 0, sigma_w = 0, a_w4 = 0, sigma_w4 = 0
    real(c_double) :: charge(MFA_NT) = 0
    real(c_double) :: alpha_c = 0
    integer(c_int32_t) :: n_kmax = 0, pad1 = 0
    integer(c_int32_t) :: s_time = 1, n_relax = 0, n_obser = 0, n_safe = 1
    real(c_double) :: dt = 0, temp = 0, x_space = 0, y_space = 0, z_space_wall = 0
    real(c_double) :: mass_type(MFA_NT) = 1
    real(c_double) :: epsil(MFA_NT, MFA_NT) = 0
    real(c_double) :: sigma_ii(MFA_NT) = 1, friction(MFA_NT) = 0
    real(c_double) :: k_bend = 0, alpha_eq = 0
    integer(c_int64_t) :: iseed = 0
    real(c_double) :: k_chain = 0, r_chain = 0
    real(c_double) :: va_spring_twall(3) = 0
    integer(c_int32_t) :: system = 3, symmetry = 0, solvent = 1, thermostat = 0
    integer(c_int32_t) :: particle_4 = 0, bending = 0
    real(c_double) :: dpd_cut_off = 0, skin = 0, lj_rc = 0, rc_coul = 0
    integer(c_int32_t) :: ewald_compat = 0, force_switch_on = 0
    integer(c_int32_t) :: dpd_vv = 0, pad2 = 0
  end type

  type, bind(C) :: mfa_params_t
    integer(c_int32_t) :: n_part, n_mon_tot
    integer(c_int32_t) :: n_mon, n_chain, n_mon_d, n_chain_d, n_mon_e, n_chain_e
    integer(c_int32_t) :: part_init_d, part_init_e
    integer(c_int32_t) :: system, symmetry, thermostat, dpd_weight
    integer(c_int32_t) :: dpd_cut_off_on, coulomb, bending, ewald_compat
    real(c_double) :: boundary(3), inv_boundary(3)
    real(c_double) :: dt, dt_2, temp
    real(c_double) :: mass_type(MFA_NT), inv_mass_type(MFA_NT)
    real(c_double) :: epsil(MFA_NT, MFA_NT), sigma(MFA_NT, MFA_NT), sigma_2(MFA_NT, MFA_NT)
    real(c_double) :: range_2(MFA_NT, MFA_NT), inv_range_2(MFA_NT, MFA_NT)
    real(c_double) :: e_shift(MFA_NT, MFA_NT), list_2(MFA_NT, MFA_NT)
    real(c_double) :: r_cut_max, r_cut_max_2, r_cut_dpd, r_cut_dpd_2
    real(c_double) :: skin, skin_2, skin_half_2
    real(c_double) :: friction1, sig
    real(c_double) :: k_chain, r_chain_2, inv_r_chain_2
    real(c_double) :: a_w, sigma_wall(MFA_NT), a_wall(MFA_NT), z_space_wall
    real(c_double) :: va_spring_twall(3)
    real(c_double) :: k_bend, alpha_eq
    integer(c_int32_t) :: n_bin(3), delta_bin(3)
    real(c_double) :: r_bin(3), inv_r_bin(3)
    real(c_double) :: charge(MFA_NT)
    real(c_double) :: alpha_c, sqrt_alpha, inv_sqrt_pi_real, rc_coul_2, v_coul_self
    integer(c_int32_t) :: n_kmax, pad0
    integer(c_int64_t) :: seed
    real(c_double) :: fso_r2_init, fso_time
    integer(c_int32_t) :: dpd_vv, pad1
  end type

  type, bind(C) :: mfa_obs_t
    real(c_double) :: v_fluid_fluid = 0, v_fluid_wall = 0, v_intra_molec = 0, v_bend = 0
    real(c_double) :: v_coul = 0, t_kin = 0, virial = 0, wall_fz_bottom = 0
    real(c_double) :: wall_fz_top = 0, z_moment = 0
    integer(c_int64_t) :: n_rebuilds = 0
  end type

  ! mfa_obs_sums: running sums of the observation stage; C r_g2_mean[3][3]
  ! ([group][axis]) appears as r_g2_mean(3,3) with r_g2_mean(axis, group)
  type, bind(C) :: mfa_obs_sums_t
    integer(c_int64_t) :: count_obs = 0
    real(c_double) :: v_total_1 = 0, v_total_2 = 0, t_total_1 = 0, t_total_2 = 0
    real(c_double) :: e_total_1 = 0, e_total_2 = 0
    real(c_double) :: v_fluid_fluid_1 = 0, v_fluid_fluid_2 = 0, t_fluid_1 = 0, t_fluid_2 = 0
    real(c_double) :: e_fluid_fluid_1 = 0, e_fluid_fluid_2 = 0, v_tot_intra_1 = 0, v_tot_intra_2 = 0
    real(c_double) :: tot_p(3) = 0, r_g2_mean(3, 3) = 0, r_par_per2(2) = 0
    integer(c_int64_t) :: count_hist = 0
    integer(c_int32_t) :: hist_dim = 0, pad = 0
  end type

  interface
    integer(c_int32_t) function mfa_abi_sizes(si, sp, so) bind(C, name="mfa_abi_sizes")
      import :: c_int32_t
      integer(c_int32_t) :: si, sp, so
    end function
    integer(c_int32_t) function mfa_derive(inp, par) bind(C, name="mfa_derive")
      import :: c_int32_t, mfa_inputs_t, mfa_params_t
      type(mfa_inputs_t), intent(in) :: inp
      type(mfa_params_t), intent(out) :: par
    end function
    integer(c_int32_t) function mfa_create(par, precision, cap, ctx) bind(C, name="mfa_create")
      import :: c_int32_t, c_ptr, mfa_params_t
      type(mfa_params_t), intent(in) :: par
      integer(c_int32_t), value :: precision, cap
      type(c_ptr) :: ctx
    end function
    integer(c_int32_t) function mfa_destroy(ctx) bind(C, name="mfa_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mfa_last_error(ctx) bind(C, name="mfa_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function mfa_set_state(ctx, r0, v, a, a_type, mic) bind(C, name="mfa_set_state")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: r0(*), v(*), a(*)
      integer(c_int32_t), intent(in) :: a_type(*), mic(*)
    end function
    integer(c_int32_t) function mfa_get_state(ctx, r0, v, a, force, mic) bind(C, name="mfa_get_state")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: r0(*), v(*), a(*), force(*)
      integer(c_int32_t) :: mic(*)
    end function
    integer(c_int32_t) function mfa_verlet_positions(ctx) bind(C, name="mfa_verlet_positions")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function mfa_check_skin(ctx, f_skin) bind(C, name="mfa_check_skin")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: f_skin
    end function
    integer(c_int32_t) function mfa_binning(ctx) bind(C, name="mfa_binning")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function mfa_zero_force(ctx) bind(C, name="mfa_zero_force")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function mfa_fluid_fluid(ctx, step, v_ff, v_coul) bind(C, name="mfa_fluid_fluid")
      import :: c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: step
      real(c_double) :: v_ff, v_coul
    end function
    integer(c_int32_t) function mfa_ewald_k(ctx, mode, v_coul) bind(C, name="mfa_ewald_k")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
      real(c_double) :: v_coul
    end function
    integer(c_int32_t) function mfa_bending(ctx, mode, v_bend) bind(C, name="mfa_bending")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
      real(c_double) :: v_bend
    end function
    integer(c_int32_t) function mfa_dipolar_correction(ctx, v_coul) bind(C, name="mfa_dipolar_correction")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: v_coul
    end function
    integer(c_int32_t) function mfa_fluid_wall(ctx, inter_type, v_fw) bind(C, name="mfa_fluid_wall")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: inter_type
      real(c_double) :: v_fw
    end function
    integer(c_int32_t) function mfa_intra_molec(ctx, v_intra) bind(C, name="mfa_intra_molec")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: v_intra
    end function
    integer(c_int32_t) function mfa_verlet_velocities(ctx, t_kin) bind(C, name="mfa_verlet_velocities")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: t_kin
    end function
    integer(c_int32_t) function mfa_new_dpd_fd(ctx, t_kin) bind(C, name="mfa_new_dpd_fd")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: t_kin
    end function
    integer(c_int32_t) function mfa_obs_init(ctx, hist_dim) bind(C, name="mfa_obs_init")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: hist_dim
    end function
    integer(c_int32_t) function mfa_observe(ctx, e, coul_in_total) bind(C, name="mfa_observe")
      import :: c_int32_t, c_ptr, mfa_obs_t
      type(c_ptr), value :: ctx
      type(mfa_obs_t) :: e
      integer(c_int32_t), value :: coul_in_total
    end function
    integer(c_int32_t) function mfa_obs_read(ctx, sums, dens_z, f_on_heads) bind(C, name="mfa_obs_read")
      import :: c_int32_t, c_ptr, c_double, mfa_obs_sums_t
      type(c_ptr), value :: ctx
      type(mfa_obs_sums_t) :: sums
      real(c_double) :: dens_z(*), f_on_heads(*)
    end function
    integer(c_int32_t) function mfa_run_steps(ctx, first_step, n, obs) bind(C, name="mfa_run_steps")
      import :: c_int32_t, c_int64_t, c_ptr, mfa_obs_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first_step
      integer(c_int32_t), value :: n
      type(mfa_obs_t) :: obs
    end function
    integer(c_int32_t) function mfa_generate(par, kind, seed, head_z, bond, min_dist, r0, v, a_type) &
        bind(C, name="mfa_generate")
      import :: c_int32_t, c_int64_t, c_double, mfa_params_t
      type(mfa_params_t), intent(in) :: par
      integer(c_int32_t), value :: kind
      integer(c_int64_t), value :: seed
      real(c_double), value :: head_z, bond, min_dist
      real(c_double) :: r0(*), v(*)
      integer(c_int32_t) :: a_type(*)
    end function
    integer(c_int32_t) function mfa_set_skin(ctx, skin) bind(C, name="mfa_set_skin")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: skin
    end function
    integer(c_int32_t) function mfa_autotune_skin(ctx, first_step, n_steps, n_cand, skins, &
        best_skin, ms_per_step) bind(C, name="mfa_autotune_skin")
      import :: c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first_step
      integer(c_int32_t), value :: n_steps, n_cand
      real(c_double) :: skins(*)
      real(c_double) :: best_skin, ms_per_step
    end function
    integer(c_int32_t) function mfa_set_cells(ctx, width) bind(C, name="mfa_set_cells")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: width
    end function
    integer(c_int32_t) function mfa_autotune_cells(ctx, first_step, n_steps, n_width, widths, &
        n_skin, skins, best_width, best_skin, ms_per_step, next_step) bind(C, name="mfa_autotune_cells")
      import :: c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: first_step
      integer(c_int32_t), value :: n_steps, n_width, n_skin
      real(c_double) :: widths(*), skins(*)
      real(c_double) :: best_width, best_skin, ms_per_step
      integer(c_int64_t) :: next_step
    end function
    integer(c_int32_t) function mfa_set_half_list(ctx, on) bind(C, name="mfa_set_half_list")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Message of the last failure of `ctx` (or of a call without a context).
  function mfa_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = mfa_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [max(n, 1)])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! The reference prints and stops on fatal conditions (e.g. init_system.f90:141-143);
  ! the driver does the same for any non-zero status.
  subroutine mfa_check(rc, what, ctx)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: what
    type(c_ptr), intent(in) :: ctx
    if (rc /= 0) then
      print '(a,a,a,i0,a,a)', " *** ", what, " failed (status ", rc, "): ", mfa_message(ctx)
      stop 1
    end if
  end subroutine

  subroutine mfa_check_abi()
    integer(c_int32_t) :: si, sp, so, rc
    type(mfa_inputs_t) :: a
    type(mfa_params_t) :: b
    type(mfa_obs_t) :: c
    rc = mfa_abi_sizes(si, sp, so)
    if (si /= c_sizeof(a) .or. sp /= c_sizeof(b) .or. so /= c_sizeof(c)) then
      print *, "*** libmfa_hip struct sizes differ from the Fortran mirror:", si, sp, so, &
        c_sizeof(a), c_sizeof(b), c_sizeof(c)
      stop 1
    end if
  end subroutine

end module mfa_hip_bind
