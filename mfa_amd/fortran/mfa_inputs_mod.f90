! Readers and writers of MFA_PROG's run-time files, for the switch set the
! library covers (WALL 2, SOLVENT 0/1, SYSTEM 2/3, optional PARTICLE_4 and
! BENDING).  The reads are list-directed like the reference's, so a file is
! interpreted the same way:
!   system_input  init_system.f90:12-144
!   mfa_input     init_params.f90:137-244
!   conf_old      init_config.f90:110-209 (formats :561-568)
!   conf_new      store_config.f90:19-61 (formats :277-284)
! The compile-time switches of control_simulation.h come from `hip_options`
! (one "key value" pair per line, '#' starts a comment).
module mfa_inputs_mod
  use iso_c_binding
  use mfa_hip_bind
  implicit none

  ! driver options (not part of the reference's files)
  character(len=16) :: run_mode = "compat"   ! compat: one call per subroutine; fused: mfa_run_steps
  integer(c_int32_t) :: precision = MFA_FP64
  integer(c_int32_t) :: list_capacity = 0
  integer :: observe = 1                     ! observation stage after n_relax (md_main.f90:159-161)
  integer :: dim_prof_dens = 200             ! its density slabs (init_obser.f90:120)

contains

  subroutine read_hip_options(inp, fname)
    type(mfa_inputs_t), intent(inout) :: inp
    character(len=*), intent(in) :: fname
    character(len=256) :: line, key
    character(len=64) :: val
    integer :: u, ios, k
    logical :: ex
    inquire(file=fname, exist=ex)
    if (.not. ex) return
    open(newunit=u, file=fname, status="old", action="read")
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      k = index(line, '#')
      if (k > 0) line = line(1:k-1)
      if (len_trim(line) == 0) cycle
      read(line, *, iostat=ios) key, val
      if (ios /= 0) cycle
      select case (trim(key))
      case ("system");          read(val, *) inp%system
      case ("symmetry");        read(val, *) inp%symmetry
      case ("solvent");         read(val, *) inp%solvent
      case ("thermostat");      read(val, *) inp%thermostat
      case ("particle_4");      read(val, *) inp%particle_4
      case ("bending");         read(val, *) inp%bending
      case ("dpd_cut_off");     read(val, *) inp%dpd_cut_off
      case ("skin");            read(val, *) inp%skin
      case ("lj_rc");           read(val, *) inp%lj_rc
      case ("rc_coul");         read(val, *) inp%rc_coul
      case ("ewald_compat");    read(val, *) inp%ewald_compat
      case ("force_switch_on"); read(val, *) inp%force_switch_on
      case ("dpd_vv");          read(val, *) inp%dpd_vv
      case ("list_capacity");   read(val, *) list_capacity
      case ("observe");         read(val, *) observe
      case ("dim_prof_dens");   read(val, *) dim_prof_dens
      case ("mode");            run_mode = trim(val)
      case ("precision")
        if (trim(val) == "mixed") then
          precision = MFA_MIXED
        else
          precision = MFA_FP64
        end if
      case default
        print *, " * hip_options: unknown key ", trim(key)
        stop 1
      end select
    end do
    close(u)
  end subroutine

  ! system_input (init_system.f90:12-144).
  subroutine read_system_input(inp)
    type(mfa_inputs_t), intent(inout) :: inp
    integer :: u, n_read, i
    real(8) :: drop_cell_x, drop_cell_y, z_skin
    open(newunit=u, file="system_input", status="old", action="read")
    do i = 1, 4
      read(u, *)
    end do
    read(u, *) inp%n_mon, inp%n_chain, inp%n_cell_w_x, inp%n_cell_w_y
    read(u, *)
    read(u, *, iostat=n_read) inp%n_mon_d, inp%n_chain_d, drop_cell_x, drop_cell_y, z_skin
    if (n_read /= 0) print *, " * Some parameter is missing in system_input file ! "
    if (inp%particle_4 /= 0) then
      read(u, *)
      read(u, *, iostat=n_read) inp%n_mon_e, inp%n_chain_e, inp%inter4
      if (n_read /= 0) then
        print *, " WARN!! missing parameter in system_input. Line for particle 4! "
        stop 1
      end if
    end if
    read(u, *)
    if (inp%particle_4 /= 0) then
      read(u, *, iostat=n_read) inp%a_w, inp%sigma_w, inp%a_w4, inp%sigma_w4
    else
      read(u, *, iostat=n_read) inp%a_w, inp%sigma_w
    end if
    if (n_read /= 0) then
      print *, " * Parameter missing in system_input! Line: a_w, sigma_w "
      stop 1
    end if
    read(u, *)
    read(u, *) inp%charge(1:4)
    read(u, *)
    read(u, *, iostat=n_read) inp%alpha_c, inp%n_kmax
    if (inp%system == 3 .and. inp%charge(1) /= 0d0) then
      print '(/a/)', "  *  For  charged brushes with implicit walls, brush' s heads MUST have q=0"
      stop 1
    end if
    close(u)
  end subroutine

  ! mfa_input (init_params.f90:137-244).
  subroutine read_mfa_input(inp)
    type(mfa_inputs_t), intent(inout) :: inp
    integer :: u, it, jt, i, f_twall(3)
    integer :: iseed32
    real(8) :: rd, k_spring_wall, k_spring_twall(3), eps_in(MFA_NT, MFA_NT)
    character(len=64) :: c_dummy
    open(newunit=u, file="mfa_input", status="old", action="read")
    read(u, *) inp%s_time
    read(u, *) inp%n_relax
    read(u, *) inp%n_obser
    read(u, *) inp%n_safe
    read(u, *) inp%dt
    read(u, *) c_dummy
    read(u, *)
    read(u, *) inp%temp
    read(u, *)
    read(u, *)
    read(u, *) inp%x_space
    read(u, *) inp%y_space
    read(u, *) inp%z_space_wall
    read(u, *)
    do it = 1, MFA_NT
      read(u, *) inp%mass_type(it)
      do jt = 1, MFA_NT
        read(u, *) eps_in(it, jt)
      end do
      read(u, *) inp%sigma_ii(it)
      read(u, *) inp%friction(it), c_dummy
    end do
    ! C layout epsil[i][j] == Fortran epsil(j,i)
    do it = 1, MFA_NT
      do jt = 1, MFA_NT
        inp%epsil(jt, it) = eps_in(it, jt)
      end do
    end do
    read(u, *) k_spring_wall, c_dummy
    if (inp%bending /= 0) then
      read(u, *) inp%k_bend
      read(u, *) inp%alpha_eq
    end if
    read(u, *)
    read(u, *) iseed32
    inp%iseed = iseed32
    read(u, *)
    read(u, *) inp%k_chain
    read(u, *) inp%r_chain
    do i = 1, 4
      read(u, *)
    end do
    read(u, *) f_twall(1:3)
    read(u, *) inp%va_spring_twall(1:3)
    read(u, *) k_spring_twall(1:3)
    close(u)
    rd = 0d0
  end subroutine

  ! conf_old (init_config.f90:110-209, formats :561-568, WALL 2).
  subroutine read_conf_old(n, a_type, r0, v, a)
    integer, intent(in) :: n
    integer(c_int32_t), intent(out) :: a_type(n)
    real(c_double), intent(out) :: r0(3, n), v(3, n), a(3, n)
    integer :: u, i, i_mon, i_chain, icx, icy, j_order
    open(newunit=u, file="conf_old", status="old", action="read")
    read(u, *)
    read(u, '(2i10)') i_mon, i_chain
    read(u, '(2i10)') icx, icy
    read(u, '(i3)') j_order
    do i = 1, n
      read(u, '(i3,3f13.7)') a_type(i), r0(1, i), r0(2, i), r0(3, i)
    end do
    do i = 1, n
      read(u, '(3f15.7)') v(1:3, i)
    end do
    do i = 1, n
      read(u, '(3f15.7)') a(1:3, i)
    end do
    close(u)
  end subroutine

  ! The full-precision sidecar conf_old.bin (a conf_new.bin renamed), if
  ! present and of n particles: r0, v, a and the wrap counters as stored, so a
  ! run resumed from it continues bit for bit.
  subroutine read_state_sidecar(n, r0, v, a, mic, found)
    integer, intent(in) :: n
    real(c_double), intent(inout) :: r0(3, n), v(3, n), a(3, n)
    integer(c_int32_t), intent(inout) :: mic(3, n)
    logical, intent(out) :: found
    integer :: u
    integer(c_int32_t) :: m
    inquire(file="conf_old.bin", exist=found)
    if (.not. found) return
    open(newunit=u, file="conf_old.bin", status="old", access="stream", form="unformatted", &
         action="read")
    read(u) m
    if (m /= n) then
      close(u)
      print *, " * conf_old.bin holds ", m, " particles, not ", n, ": ignored"
      found = .false.
      return
    end if
    read(u) r0, v, a, mic
    close(u)
  end subroutine

  ! conf_new (store_config.f90:19-61 with WALL 2), plus a full-precision
  ! sidecar conf_new.bin (stream: n, r0, v, a, mic) for exact restarts.
  subroutine write_conf_new(inp, n, a_type, r0, v, a, mic)
    type(mfa_inputs_t), intent(in) :: inp
    integer, intent(in) :: n
    integer(c_int32_t), intent(in) :: a_type(n), mic(3, n)
    real(c_double), intent(in) :: r0(3, n), v(3, n), a(3, n)
    integer :: u, i
    open(newunit=u, file="conf_new", status="replace", action="write")
    write(u, '(i1)') 1
    write(u, '(4i10)') inp%n_mon, inp%n_chain, inp%n_mon_d, inp%n_chain_d
    write(u, '(2i10)') inp%n_cell_w_x, inp%n_cell_w_y
    write(u, '(i3)') 2
    do i = 1, n
      write(u, '(i3,3f13.7)') a_type(i), r0(1, i), r0(2, i), r0(3, i)
    end do
    do i = 1, n
      write(u, '(3f15.7)') v(1:3, i)
    end do
    do i = 1, n
      write(u, '(3f15.7)') a(1:3, i)
    end do
    ! wall bookkeeping lines of store_config.f90:44-54 (all zero without walls atoms)
    write(u, '(3f15.7)') 0d0, 0d0, 0d0
    write(u, '(3f15.7)') 0d0, 0d0, 0d0
    write(u, '(3f15.7)') 0d0, 0d0, 0d0
    write(u, '(3i13)') 0, 0, 0
    write(u, '(3f15.7)') 0d0, 0d0, 0d0
    write(u, '(3i13)') 0, 0, 0
    write(u, '(3f15.7)') 0d0, 0d0, 0d0
    write(u, '(3i13)') 0, 0, 0
    close(u)
    open(newunit=u, file="conf_new.bin", status="replace", access="stream", form="unformatted")
    write(u) int(n, c_int32_t), r0, v, a, mic
    close(u)
  end subroutine

end module mfa_inputs_mod
