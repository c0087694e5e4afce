! mfa_hip_bridge.f90 -- Fortran side of the drop-in boundary.
!
! Compiled against the reference's `module commons` (mfa_common.f90) and
! linked in place of the reference objects of the step routines, these
! subroutines keep the names and argument lists md_main.f90 calls
! (md_main.f90:85-185) and forward to the C ABI of include/mfa_hip.h.
!
! The device owns the step state.  The first routine called creates the
! context from the tables init_params/init_config left in `commons` and
! uploads r0, v, a, a_type (init_config.f90:164-173).  After every
! verlet_velocities the state and energies are copied back into `commons`,
! so the reference's host consumers (observation, store_config, obser_out:
! md_main.f90:202-217) see current arrays.  Setting MFA_HIP_SYNC_EVERY=k in
! the environment copies the arrays back only every k-th step (energies every
! step).  MFA_HIP_PRECISION=mixed selects the FP32-pair path (default FP64).
!
! Bead order, types and the (3, n_part) layout are the reference's; the
! compile-time switches of control_simulation.h become the run-time fields
! of mfa_system_params, set here from the same macros.
#include "control_simulation.h"

module mfa_hip_iface
    use iso_c_binding
    implicit none

    type, bind(C) :: mfa_system_params
        integer(c_int32_t) :: n_mon, n_chain, n_mon_d, n_chain_d
        integer(c_int32_t) :: thermostat, dpd_cut_off, force_switch_on, bending, bending_melt, poiseuille
        integer(c_int32_t) :: l_box, bend_form
        real(c_double) :: boundary(3)
        real(c_double) :: epsil(16), sigma_2(16), range_2(16), e_shift(16), inv_range_2(16)
        real(c_double) :: r_cut_max, r_cut_dpd_2, skin, sig, friction, r_2_min_init, r_2_min_time
        real(c_double) :: sigma_wall(4), a_wall(4), a_w, k_chain, inv_r_chain_2, k_bend, alpha_eq, dt
        real(c_double) :: const_force(3), head_vel(2), mass_type(4)
        integer(c_int64_t) :: seed, t_idx0
        integer(c_int32_t) :: dpd_vv, reserved1
    end type

    type, bind(C) :: mfa_energies
        real(c_double) :: v_fluid_fluid, v_fluid_wall, v_intra_molec, v_bend, v_bend_melt, t_fluid, virial(6), &
                          virial_bond(6)
    end type

    interface
        integer(c_int) function mfa_hip_create(sys, n_sys, precision, ctx) bind(C, name="mfa_hip_create")
            import :: c_int, c_int32_t, c_ptr, mfa_system_params
            type(mfa_system_params), intent(in) :: sys
            integer(c_int32_t), value :: n_sys, precision
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function mfa_hip_set_state(ctx, isys, r0, v, a, a_type) bind(C, name="mfa_hip_set_state")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: isys
            real(c_double), intent(in) :: r0(*), v(*), a(*)
            integer(c_int32_t), intent(in) :: a_type(*)
        end function
        integer(c_int) function mfa_hip_get_state(ctx, isys, r0, v, a, force) bind(C, name="mfa_hip_get_state")
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: isys
            real(c_double), intent(out) :: r0(*), v(*), a(*), force(*)
        end function
        integer(c_int) function mfa_hip_set_time(ctx, t_idx) bind(C, name="mfa_hip_set_time")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: t_idx
        end function
        integer(c_int) function mfa_hip_verlet_positions(ctx) bind(C, name="mfa_hip_verlet_positions")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfa_hip_check_skin(ctx, f_skin) bind(C, name="mfa_hip_check_skin")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: f_skin
        end function
        integer(c_int) function mfa_hip_fluid_fluid(ctx) bind(C, name="mfa_hip_fluid_fluid")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfa_hip_intra_molec(ctx) bind(C, name="mfa_hip_intra_molec")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfa_hip_verlet_velocities(ctx) bind(C, name="mfa_hip_verlet_velocities")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfa_hip_fluid_wall(ctx, inter_type) bind(C, name="mfa_hip_fluid_wall")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: inter_type
        end function
        integer(c_int) function mfa_hip_bending(ctx, mode) bind(C, name="mfa_hip_bending")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: mode
        end function
        integer(c_int) function mfa_hip_bending_melt(ctx, mode) bind(C, name="mfa_hip_bending_melt")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: mode
        end function
        integer(c_int) function mfa_hip_constant_force(ctx, case_pois) bind(C, name="mfa_hip_constant_force")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: case_pois
        end function
        integer(c_int) function mfa_hip_new_dpd_fd(ctx) bind(C, name="mfa_hip_new_dpd_fd")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfa_hip_wall_wall(ctx, inter_type, sigma_w, v) bind(C, name="mfa_hip_wall_wall")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: inter_type
            real(c_double), value :: sigma_w
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function mfa_hip_get_energies(ctx, isys, e) bind(C, name="mfa_hip_get_energies")
            import :: c_int, c_ptr, c_int32_t, mfa_energies
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: isys
            type(mfa_energies), intent(out) :: e
        end function
        integer(c_int) function mfa_hip_get_f_on_heads(ctx, isys, f) bind(C, name="mfa_hip_get_f_on_heads")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: isys
            real(c_double), intent(out) :: f(*)
        end function
        type(c_ptr) function mfa_hip_last_error() bind(C, name="mfa_hip_last_error")
            import :: c_ptr
        end function
    end interface

    type(c_ptr), save :: mfa_ctx = c_null_ptr
    integer, save :: mfa_step_count = 0, mfa_sync_every = 1

contains

    subroutine mfa_chk(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: k
        if (rc == 0) return
        call c_f_pointer(mfa_hip_last_error(), msg, [512])
        k = 0
        do while (k < 512)
            if (msg(k + 1) == c_null_char) exit
            k = k + 1
        end do
        print *, "[mfa_hip] ", what, " failed: ", msg(1:k)
        stop 1
    end subroutine mfa_chk

end module mfa_hip_iface


module mfa_hip_state
    use iso_c_binding
    use mfa_hip_iface
    implicit none
contains

    ! Create the device context from `commons` on first use.
    subroutine mfa_bridge_ensure()
        use commons
        type(mfa_system_params) :: p
        integer :: i, j, k
        character(len=32) :: env
        integer :: ln, st
        integer(c_int32_t) :: prec
        if (c_associated(mfa_ctx)) return
        call get_environment_variable("MFA_HIP_SYNC_EVERY", env, ln, st)
        if (st == 0 .and. ln > 0) read(env(1:ln), *) mfa_sync_every
        mfa_sync_every = max(1, mfa_sync_every)
        prec = 0_c_int32_t                   ! FP64; MFA_HIP_PRECISION=mixed selects the mixed path
        call get_environment_variable("MFA_HIP_PRECISION", env, ln, st)
        if (st == 0 .and. ln >= 5) then
            if (env(1:5) == "mixed") prec = 1_c_int32_t
        end if
        p%n_mon = n_mon; p%n_chain = n_chain; p%n_mon_d = n_mon_d; p%n_chain_d = n_chain_d
#if THERMOSTAT == 0
        p%thermostat = 1
#elif defined(MFA_HIP_LANGEVIN)
        ! lgv_forces.f90, per bead, as fluid_fluid_old.f90:233-236 calls it
        p%thermostat = 2
#else
        ! fluid_fluid.f90:455-458 leaves the lgv_forces call commented out
        p%thermostat = 0
#endif
#ifdef DPD_CUT_OFF
        p%dpd_cut_off = 1
#else
        p%dpd_cut_off = 0
#endif
#ifdef FORCE_SWITCH_ON
        p%force_switch_on = 1
#else
        p%force_switch_on = 0
#endif
#ifdef BENDING
        p%bending = 1
#else
        p%bending = 0
#endif
#ifdef BENDING_MELT
        p%bending_melt = 1
#else
        p%bending_melt = 0
#endif
#ifdef POISEUILLE
        p%poiseuille = 1
#else
        p%poiseuille = 0
#endif
#ifdef L_BOX
        p%l_box = L_BOX
#else
        p%l_box = 1                          ! headers without L_BOX (Verlet-list builds): cells of one cutoff
#endif
        p%bend_form = 0
#ifdef DPD_VV
        p%dpd_vv = 1
#else
        p%dpd_vv = 0
#endif
        p%reserved1 = 0
        p%boundary = boundary
        k = 0
        do j = 1, 4
            do i = 1, 4
                k = k + 1
                p%epsil(k) = epsil(i, j); p%sigma_2(k) = sigma_2(i, j); p%range_2(k) = range_2(i, j)
                p%e_shift(k) = e_shift(i, j); p%inv_range_2(k) = inv_range_2(i, j)
            end do
        end do
        p%r_cut_max = r_cut_max
#ifdef DPD_CUT_OFF
        p%r_cut_dpd_2 = r_cut_dpd_2
#else
        p%r_cut_dpd_2 = 0.0d0
#endif
#if BIN_TYPE == 0 || BIN_TYPE == 1
        p%skin = skin                        ! init_params.f90:823-835; rebuild past skin/2 (check_skin.f90:21)
#else
        p%skin = 0.0d0                       ! BIN_TYPE 2: cells rebuilt every step
#endif
        p%sig = sig; p%friction = friction(1)
        p%r_2_min_init = r_2_min_init; p%r_2_min_time = r_2_min_time
        p%sigma_wall = sigma_wall(1:4); p%a_wall = a_wall(1:4); p%a_w = a_w
        p%k_chain = k_chain; p%inv_r_chain_2 = inv_r_chain_2
        p%k_bend = k_bend; p%alpha_eq = alpha_eq
        p%dt = dt
        p%const_force = const_force
        p%head_vel = va_spring_twall(1:2)
        p%mass_type = mass_type(1:4)
        p%seed = int(iseed, c_int64_t)
        p%t_idx0 = int(s_time, c_int64_t)
        call mfa_chk(mfa_hip_create(p, 1_c_int32_t, prec, mfa_ctx), "mfa_hip_create")
        call mfa_chk(mfa_hip_set_state(mfa_ctx, 0_c_int32_t, r0, v, a, a_type), "mfa_hip_set_state")
    end subroutine mfa_bridge_ensure

    subroutine mfa_bridge_pull(arrays)
        use commons
        logical, intent(in) :: arrays
        type(mfa_energies) :: e
        call mfa_chk(mfa_hip_get_energies(mfa_ctx, 0_c_int32_t, e), "mfa_hip_get_energies")
        v_fluid_fluid = e%v_fluid_fluid
        v_fluid_wall = e%v_fluid_wall
        v_intra_molec = e%v_intra_molec
        v_bend = e%v_bend
        v_bend_melt = e%v_bend_melt
        t_fluid = e%t_fluid
        if (arrays) then
            call mfa_chk(mfa_hip_get_state(mfa_ctx, 0_c_int32_t, r0, v, a, force), "mfa_hip_get_state")
            if (n_chain > 0) call mfa_chk(mfa_hip_get_f_on_heads(mfa_ctx, 0_c_int32_t, f_on_heads), &
                                          "mfa_hip_get_f_on_heads")
        end if
    end subroutine mfa_bridge_pull

end module mfa_hip_state


! ---- the step routines, same names and arguments as the reference --------

subroutine verlet_positions()                     ! verlet_positions.f90:1
    use commons, only: s_time, i_time
    use mfa_hip_state
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_set_time(mfa_ctx, int(s_time + i_time - 1, c_int64_t)), "mfa_hip_set_time")
    call mfa_chk(mfa_hip_verlet_positions(mfa_ctx), "verlet_positions")
end subroutine verlet_positions

! BIN_TYPE 0 / 1 drivers call check_skin after verlet_positions and, when
! f_skin is 1, binning or my_binning (md_main.f90:92-102; also once at the
! start, init_config.f90:255-257, conf_default.f90:202-204).  The device runs
! the skin test and the rebuild inside verlet_positions and keeps its own
! lists, so check_skin reports the device's f_skin and the two binning
! routines only reset it (binning.f90:125, my_binning.f90:130); the host
! lists they filled (ff_list, r0_old, mic_old) were read only by the replaced
! fluid_fluid, new_dpd_fd and check_skin.
subroutine check_skin()                           ! check_skin.f90:4
    use commons, only: f_skin
    use mfa_hip_state
    integer(c_int32_t) :: fs
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_check_skin(mfa_ctx, fs), "check_skin")
    f_skin = fs
end subroutine check_skin

subroutine binning()                              ! binning.f90:2
    use commons, only: f_skin
    f_skin = 0
end subroutine binning

subroutine my_binning()                           ! my_binning.f90:5
    use commons, only: f_skin
    f_skin = 0
end subroutine my_binning

subroutine fluid_fluid()                          ! fluid_fluid.f90:1 (includes md_main.f90:107 force=0)
    use mfa_hip_state
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_fluid_fluid(mfa_ctx), "fluid_fluid")
end subroutine fluid_fluid

subroutine fluid_wall(inter_type)                 ! fluid_wall.f90:2
    use mfa_hip_state
    integer, intent(in) :: inter_type
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_fluid_wall(mfa_ctx, int(inter_type, c_int32_t)), "fluid_wall")
end subroutine fluid_wall

subroutine wall_wall(inter_type)                  ! wall_wall.f90:1
    use commons, only: v_wall_wall, sigma_w
    use mfa_hip_state
    integer, intent(in) :: inter_type
    real(c_double) :: vw
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_wall_wall(mfa_ctx, int(inter_type, c_int32_t), sigma_w, vw), "wall_wall")
    v_wall_wall = vw
end subroutine wall_wall

subroutine intra_molec()                          ! intra_molec.f90:1
    use mfa_hip_state
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_intra_molec(mfa_ctx), "intra_molec")
end subroutine intra_molec

subroutine bending(mode_bend)                     ! bending.f90:2
    use mfa_hip_state
    integer, intent(in) :: mode_bend
    if (mode_bend == 0) return
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_bending(mfa_ctx, int(mode_bend, c_int32_t)), "bending")
end subroutine bending

subroutine bending_melt(mode_bend)                ! bending_melt.f90
    use mfa_hip_state
    integer, intent(in) :: mode_bend
    if (mode_bend == 0) return
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_bending_melt(mfa_ctx, int(mode_bend, c_int32_t)), "bending_melt")
end subroutine bending_melt

subroutine constant_force(case_pois)              ! constant_force.f90:1
    use mfa_hip_state
    integer, intent(in) :: case_pois
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_constant_force(mfa_ctx, int(case_pois, c_int32_t)), "constant_force")
end subroutine constant_force

subroutine verlet_velocities()                    ! verlet_velocities.f90:1 (+ energy bookkeeping :132-155)
    use commons, only: v_total, v_wall_wall, v_fluid_wall, v_fluid_fluid, v_intra_molec, v_bend, v_bend_melt, &
                       t_fluid, t_total, t_wall, e_total
    use mfa_hip_state
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_verlet_velocities(mfa_ctx), "verlet_velocities")
    mfa_step_count = mfa_step_count + 1
    call mfa_bridge_pull(mod(mfa_step_count, mfa_sync_every) == 0)
    v_total = v_wall_wall + v_fluid_wall + v_fluid_fluid + v_intra_molec
#ifdef BENDING
    v_total = v_total + v_bend
#endif
#ifdef BENDING_MELT
    v_total = v_total + v_bend_melt
#endif
    t_total = t_wall + t_fluid
    e_total = v_total + t_total
end subroutine verlet_velocities

subroutine new_dpd_fd()                           ! new_dpd_fd.f90:1 (md_main.f90:198, DPD_VV)
    use mfa_hip_state
    call mfa_bridge_ensure()
    call mfa_chk(mfa_hip_new_dpd_fd(mfa_ctx), "new_dpd_fd")
    call mfa_bridge_pull(mod(mfa_step_count, mfa_sync_every) == 0)    ! force and a changed again
end subroutine new_dpd_fd
