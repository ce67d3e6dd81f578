! Test infrastructure, not product code.  A driver of the reference's own
! hot-path subroutines (init_system, init_params, init_config, binning,
! check_skin, verlet_positions, fluid_fluid, fluid_wall, intra_molec,
! verlet_velocities), linked against objects that build_ref.sh compiles from
! the reference's Fortran sources into oracle/_ref/.  The tests use it only
! to pin the C oracle against the real reference; nothing ships it.
!
! Run in a directory holding system_input, mfa_input and conf_old:
!   ref_md <force_step> <n_steps> [trace | snap <every>]
! writes
!   ref_init.bin   the state, the cell of every particle, the pair list of
!                  the first binning, and one force pass at r_time =
!                  force_step*dt (md_main.f90:84-132 without integrating);
!   ref_final.bin  the state after n_steps of the main loop
!                  (md_main.f90:68-147, observation and storage left out).
! Record layout (stream, native endianness): int32 n_part, int32 n_rebuilds,
! float64 v_fluid_fluid, v_fluid_wall, v_intra_molec, t_fluid, v_coul, v_bend, then
! r0, v, a, force as float64 (3, n_part) and mic_count as int32 (3, n_part);
! ref_init.bin adds the cells int32 (3, n_part), int32 n_pairs and the pairs
! int32 (2, n_pairs), each pair as (min, max).
! With `trace`, ref_trace.bin gets per step (after verlet_velocities) two
! float64: the sum of z over type-2 beads and the sum over all particles of
! the top wall's z force (the second half of fluid_wall.f90:107-121).
! Test infrastructure, used by the adjoint's finite-difference check.
! With `snap <every>`, ref_snap.bin gets r0 and v as float64 (3, n_part)
! after every <every>-th step (the statistics fixtures of the DPD runs).
! With `obs`, the reference's own observation stage runs as md_main.f90
! does it: init_obser before the loop, observation after every step past
! n_relax (md_main.f90:159-161), obser_out at the end; its files (fort.40-43,
! fort.80 and, built with PROFILES, dens_prof.mide) are what the tests read.
! With `store`, the reference's restart writer runs after the loop
! (store_config(2), md_main.f90:168-169) and leaves conf_new.
program ref_md
#include "control_simulation.h"
  use commons
  implicit none
  character(len=32) :: arg
  integer :: force_step, n_steps, n_rebuilds, n_pairs, k, j, u, ut
  integer, allocatable :: cells(:,:)
  logical :: trace, snap, obs, store
  integer :: snap_every, us
  real(kind=8) :: zsum, ftop, izt

  call get_command_argument(1, arg)
  read(arg, *) force_step
  call get_command_argument(2, arg)
  read(arg, *) n_steps
  trace = .false.
  snap = .false.
  obs = .false.
  store = .false.
  snap_every = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    if (trim(arg) == "obs") then
      obs = .true.
    else if (trim(arg) == "store") then
      store = .true.
    else if (trim(arg) == "snap") then
      snap = .true.
      call get_command_argument(4, arg)
      read(arg, *) snap_every
      open(newunit=us, file="ref_snap.bin", access="stream", form="unformatted", status="replace")
    else
      trace = .true.
      open(newunit=ut, file="ref_trace.bin", access="stream", form="unformatted", status="replace")
    end if
  end if

  call init_system()
  call init_params()
  call init_config()      ! ends with a binning (init_config.f90:251-255)

  ! cells of the binning just made, by binning.f90:27-29's formula
  allocate(cells(3, n_part))
  do k = 1, n_part
    cells(1, k) = mod(int(r0(1, k) * inv_r_bin_x), n_bin_x)
    cells(2, k) = mod(int(r0(2, k) * inv_r_bin_y), n_bin_y)
    cells(3, k) = mod(int(r0(3, k) * inv_r_bin_z), n_bin_z)
  end do
  n_pairs = 0
  do k = 1, n_part
    n_pairs = n_pairs + ff_list(0, k)
  end do

  i_time = 0
  r_time = dble(force_step) * dt
  force(:,:) = 0.0
  call fluid_fluid()
#if SYMMETRY == 0
  call fluid_wall(wall_flag)
#endif
  call intra_molec()
  call angles()
  call coulomb_k()
  open(newunit=u, file="ref_init.bin", access="stream", form="unformatted", status="replace")
  call put_state(u, 0)
  write(u) cells
  write(u) n_pairs
  do k = 1, n_part
    do j = 1, ff_list(0, k)
      write(u) min(k, ff_list(j, k)), max(k, ff_list(j, k))
    end do
  end do
  close(u)

  ! main loop, md_main.f90:68-161
  if (obs) call init_obser()
  n_rebuilds = 0
  do i_time = 1, n_steps
    r_time = dble(s_time + i_time - 1) * dt
    call verlet_positions()
    call check_skin
    if (f_skin .eq. 1) then
      call binning
      n_rebuilds = n_rebuilds + 1
    end if
    force(:,:) = 0.0
    call fluid_fluid()
#if SYMMETRY == 0
    call fluid_wall(wall_flag)
#endif
    call intra_molec
    call angles()
    call coulomb_k()
    call verlet_velocities()
#ifdef DPD_VV
    call new_dpd_fd()
#endif
    if (trace) then
      zsum = 0.d0
      ftop = 0.d0
      do k = 1, n_part
        if (a_type(k) == 2) zsum = zsum + r0(3, k)
#if SYMMETRY == 0
        izt = 1.d0 / (z_space_wall - r0(3, k))
        ftop = ftop - 9.d0 * abs(a_wall(a_type(k))) * sigma_wall(a_type(k))**9 * izt**10 &
                    + 3.d0 * a_wall(a_type(k)) * sigma_wall(a_type(k))**3 * izt**4
#endif
      end do
      write(ut) zsum, ftop
    end if
    if (snap) then
      if (mod(i_time, snap_every) == 0) write(us) r0(:, 1:n_part), v(:, 1:n_part)
    end if
    if (obs .and. i_time > n_relax) call observation
  end do
  if (obs) call obser_out()
  if (store) call store_config(2)
  if (trace) close(ut)
  if (snap) close(us)
  open(newunit=u, file="ref_final.bin", access="stream", form="unformatted", status="replace")
  call put_state(u, n_rebuilds)
  close(u)

contains

  ! chain stiffness (md_main.f90:108-111)
  subroutine angles()
#ifdef BENDING
    call bending(2)
#endif
  end subroutine angles

  ! k-space Ewald and the slab correction of charged systems (md_main.f90:128-132)
  subroutine coulomb_k()
#if SYSTEM == 2 || SYSTEM == 3
    call ewald_k(1)
#  if SYMMETRY == 0
    call dipolar_correction()
#  endif
#endif
  end subroutine coulomb_k

  subroutine put_state(unit, nreb)
    integer, intent(in) :: unit, nreb
    real(kind=8) :: vb
    vb = 0.d0
#ifdef BENDING
    vb = v_bend
#endif
    write(unit) n_part, nreb
    write(unit) v_fluid_fluid, v_fluid_wall, v_intra_molec, t_fluid, v_coul, vb
    write(unit) r0(:, 1:n_part), v(:, 1:n_part), a(:, 1:n_part), force(:, 1:n_part)
    write(unit) mic_count(:, 1:n_part)
  end subroutine put_state

end program ref_md
