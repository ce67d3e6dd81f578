! Driver of the force-and-integrate loop on MI355X: md_main.f90's structure
! (init, main time loop md_main.f90:68-184, final output) with every hot-path
! subroutine executed by libmfa_hip through iso_c_binding.
!
! Run in a directory holding system_input, mfa_input, hip_options and
! (optionally) conf_old.  Writes fort.60 every 10 steps
! (verlet_velocities.f90:206-227), conf_new (+ conf_new.bin) every n_safe
! steps and at the end, and the wall time.  After n_relax steps the
! observation stage runs on the device every step (md_main.f90:159-161):
! fort.40-43 and fort.80 every 5 observed steps (observation.f90:295-330)
! and dens_prof.mide at the end (functions.f90:131-160, histo mode 1).
module md_hip_state
  use iso_c_binding
  use mfa_hip_bind
  implicit none
  type(c_ptr) :: ctx = c_null_ptr
  type(mfa_inputs_t) :: inp
  type(mfa_params_t) :: par
  integer :: n_part
  real(c_double), allocatable :: r0(:,:), v(:,:), a(:,:), force(:,:)
  integer(c_int32_t), allocatable :: a_type(:), mic_count(:,:)
  integer(c_int32_t) :: f_skin
  real(c_double) :: v_fluid_fluid, v_coul, v_fluid_wall, v_intra_molec, t_fluid, v_bend

contains

  ! Thin wrappers with the reference's names (md_main.f90:73-147).
  subroutine verlet_positions()
    call mfa_check(mfa_verlet_positions(ctx), "verlet_positions", ctx)
  end subroutine
  subroutine check_skin()
    call mfa_check(mfa_check_skin(ctx, f_skin), "check_skin", ctx)
  end subroutine
  subroutine binning()
    call mfa_check(mfa_binning(ctx), "binning", ctx)
    f_skin = 0
  end subroutine
  subroutine zero_force()
    call mfa_check(mfa_zero_force(ctx), "force = 0", ctx)
  end subroutine
  subroutine fluid_fluid(step)
    integer(c_int64_t), intent(in) :: step
    call mfa_check(mfa_fluid_fluid(ctx, step, v_fluid_fluid, v_coul), "fluid_fluid", ctx)
  end subroutine
  subroutine fluid_wall(inter_type)
    integer, intent(in) :: inter_type
    call mfa_check(mfa_fluid_wall(ctx, int(inter_type, c_int32_t), v_fluid_wall), "fluid_wall", ctx)
  end subroutine
  subroutine intra_molec()
    call mfa_check(mfa_intra_molec(ctx, v_intra_molec), "intra_molec", ctx)
  end subroutine
  subroutine bending(mode)
    integer, intent(in) :: mode
    call mfa_check(mfa_bending(ctx, int(mode, c_int32_t), v_bend), "bending", ctx)
  end subroutine
  subroutine ewald_k(mode)
    integer, intent(in) :: mode
    real(c_double) :: dv
    call mfa_check(mfa_ewald_k(ctx, int(mode, c_int32_t), dv), "ewald_k", ctx)
    v_coul = v_coul + dv
  end subroutine
  subroutine dipolar_correction()
    real(c_double) :: dv
    call mfa_check(mfa_dipolar_correction(ctx, dv), "dipolar_correction", ctx)
    v_coul = v_coul + dv
  end subroutine
  subroutine verlet_velocities()
    call mfa_check(mfa_verlet_velocities(ctx, t_fluid), "verlet_velocities", ctx)
  end subroutine
  subroutine new_dpd_fd()
    call mfa_check(mfa_new_dpd_fd(ctx, t_fluid), "new_dpd_fd", ctx)
  end subroutine

  subroutine fetch_state()
    call mfa_check(mfa_get_state(ctx, r0, v, a, force, mic_count), "get_state", ctx)
  end subroutine

end module md_hip_state

program md_pb_hip
  use iso_c_binding
  use mfa_hip_bind
  use mfa_inputs_mod
  use md_hip_state
  implicit none
  integer :: i_time, tot_time, i, u60, nchunk, u40, u41, u42, u43, u80
  integer(c_int64_t) :: step
  integer(c_int32_t) :: rc
  logical :: have_conf, exact_restart
  real(8) :: inv_N, v_total, e_total, vcm(3), mtot
  integer(8) :: c0, c1, crate
  type(mfa_obs_t) :: obs
  type(mfa_obs_sums_t) :: osum
  real(c_double), allocatable :: dens_z(:,:), f_on_heads(:,:)

  call system_clock(c0, crate)
  print '(/a/)', " *** MFA_PROG on MI355X (libmfa_hip) ***"
  call mfa_check_abi()

  ! ---- initialisation (md_main.f90:51-53)
  call read_hip_options(inp, "hip_options")
  call read_system_input(inp)
  call read_mfa_input(inp)
  rc = mfa_derive(inp, par)
  call mfa_check(rc, "init_params (derived constants)", c_null_ptr)
  n_part = par%n_part
  allocate(r0(3, n_part), v(3, n_part), a(3, n_part), force(3, n_part))
  allocate(a_type(n_part), mic_count(3, n_part))
  mic_count = 0
  inquire(file="conf_old", exist=have_conf)
  exact_restart = .false.
  if (have_conf) then
    call read_conf_old(n_part, a_type, r0, v, a)
    call read_state_sidecar(n_part, r0, v, a, mic_count, exact_restart)
    if (exact_restart) print '(a)', " * conf_old.bin: full-precision state and wrap counters restored"
  else
    print '(/a/)', " *   'conf_old' file not found. Using the seeded generator ..."
    rc = mfa_generate(par, merge(1_c_int32_t, 2_c_int32_t, inp%symmetry == 0), inp%iseed, &
                      1d0, 0.97d0, 0.8d0, r0, v, a_type)
    call mfa_check(rc, "generate", c_null_ptr)
    a = 0d0
  end if
  ! SYSTEM 2/3 in bulk: remove the centre-of-mass velocity (init_config.f90:336-356);
  ! an exact restart keeps the stored velocities
  if (inp%symmetry == 1 .and. .not. exact_restart) then
    vcm = 0d0
    mtot = 0d0
    do i = 1, n_part
      vcm(:) = vcm(:) + par%mass_type(a_type(i)) * v(:, i)
      mtot = mtot + par%mass_type(a_type(i))
    end do
    vcm(:) = vcm(:) / mtot
    do i = 1, n_part
      v(1, i) = v(1, i) - vcm(1)
      v(2, i) = v(2, i) - vcm(2)
      v(3, i) = v(3, i) - vcm(3)
    end do
  end if
  rc = mfa_create(par, precision, list_capacity, ctx)
  call mfa_check(rc, "create", c_null_ptr)
  call mfa_check(mfa_set_state(ctx, r0, v, a, a_type, mic_count), "set_state", ctx)
  call binning()          ! init_config.f90:251-255

  if (inp%system == 3 .and. inp%symmetry == 0) then
    inv_N = 1d0 / dble(par%n_mon_tot - par%n_chain)   ! init_params.f90:481-486
  else
    inv_N = 1d0 / dble(par%n_mon_tot)
  end if
  open(newunit=u60, file="fort.60", status="replace", action="write")
  if (observe /= 0) then   ! init_obser.f90:15-91
    call mfa_check(mfa_obs_init(ctx, int(dim_prof_dens, c_int32_t)), "init_obser", ctx)
    allocate(dens_z(dim_prof_dens, 3), f_on_heads(3, max(par%n_chain, 1)))
    open(newunit=u40, file="fort.40", status="replace", action="write")
    open(newunit=u41, file="fort.41", status="replace", action="write")
    open(newunit=u42, file="fort.42", status="replace", action="write")
    open(newunit=u43, file="fort.43", status="replace", action="write")
    open(newunit=u80, file="fort.80", status="replace", action="write")
  end if

  ! ---- main time loop (md_main.f90:68-184)
  tot_time = inp%n_relax + inp%n_obser
  if (trim(run_mode) == "fused") then
    i_time = 0
    do while (i_time < tot_time)
      ! blocks of 10 steps; one step at a time while observing (every step
      ! past n_relax), so that each observed step is a block's last
      nchunk = min(10 - mod(i_time, 10), tot_time - i_time)
      if (observe /= 0 .and. i_time >= inp%n_relax) nchunk = 1
      if (observe /= 0 .and. i_time < inp%n_relax) nchunk = min(nchunk, inp%n_relax - i_time)
      step = int(inp%s_time + i_time, c_int64_t)
      call mfa_check(mfa_run_steps(ctx, step, int(nchunk, c_int32_t), obs), "run_steps", ctx)
      i_time = i_time + nchunk
      if (mod(i_time, 10) == 0) call write_fort60(i_time, obs%v_fluid_wall, obs%v_fluid_fluid, &
                                                  obs%v_intra_molec, obs%t_kin, obs%v_coul)
      if (observe /= 0 .and. i_time > inp%n_relax) call observation(i_time, obs)
      if (mod(i_time, inp%n_safe) == 0) call store(.true.)
    end do
  else
    do i_time = 1, tot_time
      step = int(inp%s_time + i_time - 1, c_int64_t)
      call verlet_positions()
      call check_skin()
      if (f_skin == 1) call binning()
      call zero_force()
      call fluid_fluid(step)
      if (inp%symmetry == 0) call fluid_wall(2)
      call intra_molec()
      if (inp%bending /= 0) call bending(2)            ! md_main.f90:108-111
      if (inp%system == 2 .or. inp%system == 3) then   ! md_main.f90:128-132
        call ewald_k(1)
        if (inp%symmetry == 0) call dipolar_correction()
      end if
      call verlet_velocities()
      if (inp%dpd_vv /= 0) call new_dpd_fd()           ! md_main.f90:149-155
      if (mod(i_time, 10) == 0) call write_fort60(i_time, v_fluid_wall, v_fluid_fluid, &
                                                   v_intra_molec, t_fluid, v_coul)
      if (observe /= 0 .and. i_time > inp%n_relax) then
        obs%v_fluid_fluid = v_fluid_fluid
        obs%v_fluid_wall = v_fluid_wall
        obs%v_intra_molec = v_intra_molec
        obs%t_kin = t_fluid
        obs%v_coul = v_coul
        call observation(i_time, obs)
      end if
      if (mod(i_time, inp%n_safe) == 0) call store(.true.)
    end do
  end if
  call store(.false.)
  close(u60)
  if (observe /= 0) call obser_out()
  rc = mfa_destroy(ctx)
  call system_clock(c1)
  print *, ' WALL TIME (s)= ', dble(c1 - c0) / dble(crate)

contains

  ! observation.f90:17-330 for the step just made (energies in e): the sums
  ! and profiles on the device; every 5 observed steps the lines of fort.40,
  ! 41, 42, 43 and 80.  v_total carries v_coul on the steps where the
  ! reference's verlet_velocities has just added it (every tenth call,
  ! verlet_velocities.f90:199-222).
  subroutine observation(it, e)
    integer, intent(in) :: it
    type(mfa_obs_t), intent(inout) :: e
    real(8) :: inv_count_obs, r_time
    integer :: k
    call mfa_check(mfa_observe(ctx, e, merge(1_c_int32_t, 0_c_int32_t, &
                   (inp%system == 2 .or. inp%system == 3) .and. mod(it, 10) == 0)), "observation", ctx)
    if (mod(it - inp%n_relax, 5) /= 0) return
    call mfa_check(mfa_obs_read(ctx, osum, dens_z, f_on_heads), "observation", ctx)
    inv_count_obs = 1d0 / dble(osum%count_obs)
    write(u40, '(i7,6f15.5)') it, osum%v_total_1 * inv_count_obs, osum%t_total_1 * inv_count_obs, &
        osum%e_total_1 * inv_count_obs, osum%tot_p(:) * inv_count_obs
    write(u41, '(i7,3f15.5)') it, e%v_fluid_fluid, e%v_intra_molec, e%v_fluid_wall
    write(u42, '(i7,8f15.7)') it, osum%r_g2_mean(:, 1) * inv_count_obs, &
        sum(osum%r_g2_mean(:, 1)) * inv_count_obs, osum%r_g2_mean(:, 2) * inv_count_obs, &
        sum(osum%r_g2_mean(:, 2)) * inv_count_obs
    write(u43, '(i7,2f15.7)') it, osum%r_par_per2(1:2) * inv_count_obs
    ! -sum(f_on_heads(:, heads), dim=1): per head, the sum of its force
    ! components, upper-wall chains first (observation.f90:326-328; WALL 2)
    r_time = dble(inp%s_time + it - 1) * inp%dt
    write(u80, '(7e18.7)') r_time, (-sum(f_on_heads(:, k)), k = 1, par%n_chain / 2), &
        (-sum(f_on_heads(:, k)), k = par%n_chain / 2 + 1, par%n_chain)
  end subroutine

  ! obser_out's density profile (functions.f90:131-160, histo mode 1):
  ! dens_prof.mide, slab centres and the counts per unit volume and observation.
  subroutine obser_out()
    real(8) :: surf, r_dummy, r_bin
    integer :: u73, k
    call mfa_check(mfa_obs_read(ctx, osum, dens_z, f_on_heads), "obser_out", ctx)
    close(u40)
    close(u41)
    close(u42)
    close(u43)
    close(u80)
    if (osum%count_hist == 0) return
    surf = par%boundary(1) * par%boundary(2)
    dens_z(:, :) = dens_z(:, :) * dble(dim_prof_dens) * par%inv_boundary(3) / (dble(osum%count_hist) * surf)
    open(newunit=u73, file="dens_prof.mide", status="replace", action="write")
    r_dummy = par%boundary(3) / dble(dim_prof_dens)
    r_bin = r_dummy / 2.
    do k = 1, dim_prof_dens
      write(u73, '(4f16.7)') r_bin, dens_z(k, 1), dens_z(k, 2), dens_z(k, 3)
      r_bin = r_bin + r_dummy
    end do
    close(u73)
  end subroutine

  ! fort.60 line for SYSTEM 2/3 (verlet_velocities.f90:185-227).
  subroutine write_fort60(it, vfw, vff, vin, tkin, vc)
    integer, intent(in) :: it
    real(c_double), intent(in) :: vfw, vff, vin, tkin, vc
    v_total = vfw + vff + vin
    e_total = v_total + tkin
    v_total = v_total + vc
    write(u60, '(i7,4g17.5)') it, v_total * inv_N, e_total * inv_N, tkin * inv_N, vc * inv_N
  end subroutine

  ! store_config (md_main.f90:168): conf_new and its full-precision sidecar.
  ! Inside the loop the list is rebuilt right after, as a run resumed from
  ! these files rebuilds it at its start: the mixed form's FP32 view is taken
  ! relative to the cells of the last rebuild, so both runs then repeat the
  ! same arithmetic and the resumed run continues bit for bit.
  subroutine store(rebin)
    logical, intent(in) :: rebin
    call fetch_state()
    call write_conf_new(inp, n_part, a_type, r0, v, a, mic_count)
    if (rebin) call binning()
  end subroutine

end program md_pb_hip
