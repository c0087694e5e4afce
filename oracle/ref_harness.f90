! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! bind(C) entry points around the reference's own step routines (compiled
! from /root/reference by oracle/ref.mk into oracle/_ref/), so the tests can
! run the real fluid_fluid / intra_molec / fluid_wall / bending /
! constant_force / verlet_* / check_fluid_fluid on a given state and compare
! the CPU restatement (oracle/mfa_oracle.c) with them, and so bench.py can time
! the reference's OpenMP step on the host cores.  Module state is set here
! directly (the role init_system/init_params/init_config play in the
! reference, init_config.f90:11-22,258-391); nothing is read from files.
module ref_harness
    use iso_c_binding
    use commons
    use util, only: check_fluid_fluid
    use ziggurat, only: zigset
#ifdef _OPENMP
    use omp_lib
    use Par_Zig_mod, only: par_zigset
#endif
    implicit none
contains

    subroutine ref_setup(nm, nc, nmd, ncd, bnd, eps, sg2, rg2, shf, irg2, rcmax, sg, fric, sw, aw, a_w_, &
                         kch, irch2, kb, aeq, dt_, cf, mtype, seed) bind(C, name="ref_setup")
        integer(c_int), value :: nm, nc, nmd, ncd, seed
        real(c_double), intent(in) :: bnd(3), eps(16), sg2(16), rg2(16), shf(16), irg2(16), sw(4), aw(4), cf(3), mtype(4)
        real(c_double), value :: rcmax, sg, fric, a_w_, kch, irch2, kb, aeq, dt_
        integer :: i, j, k, nt
        integer, allocatable :: seeds(:)
        n_mon = nm; n_chain = nc; n_mon_d = nmd; n_chain_d = ncd
        n_mon_tot = nm * nc + nmd * ncd
        n_part = n_mon_tot
        part_init_d = nm * nc
        part_init_e = n_mon_tot
        boundary = bnd
        inv_boundary = 1.0d0 / bnd
        z_space_wall = bnd(3)
        k = 0
        do j = 1, 4
            do i = 1, 4
                k = k + 1
                epsil(i, j) = eps(k); sigma_2(i, j) = sg2(k); range_2(i, j) = rg2(k)
                e_shift(i, j) = shf(k); inv_range_2(i, j) = irg2(k)
            end do
        end do
        r_cut_max = rcmax
        sig = sg
        friction = fric
        sigma_wall = sw; a_wall = aw; a_w = a_w_
        k_chain = kch; inv_r_chain_2 = irch2; r_chain_2 = 1.0d0 / irch2
        k_bend = kb; alpha_eq = aeq
        k_or = 0.0d0; alpha_or = 0.0d0; v_or = 0.0d0
        dt = dt_; dt_2 = dt_**2
        r_2_min = 0.0d0; r_2_min_init = 0.0d0; r_2_min_time = 0.0d0; r_time = 0.0d0
        const_force = cf
        mass_type = mtype
        va_spring_twall = 0.0d0; f_twall = 0; r0_twall = 0.0d0
        v_wall_wall = 0.0d0; t_wall = 0.0d0; v_intra_molec_e = 0.0d0
        time_ave_count_2 = -huge(1)      ! never write fort.60/61
        inv_N = 1.0d0 / dble(max(1, n_mon_tot - n_chain))
        if (allocated(r0)) deallocate(r0, v, a, force, a_type, mass, inv_mass, ext_force, f_on_heads)
        allocate(r0(3, n_part), v(3, n_part), a(3, n_part), force(3, n_part), a_type(n_part))
        allocate(mass(n_part), inv_mass(n_part), ext_force(3, n_part), f_on_heads(3, max(1, n_chain)))
        if (allocated(n_cells)) deallocate(n_cells, l_cell, inv_l_cell, r_nei, l_nei, part_in_cell, cell_neigh_ls)
        allocate(n_cells(3), l_cell(3), inv_l_cell(3), r_nei(n_part), l_nei(n_part))
        call make_binning(1, 3, boundary, r_cut_max, l_cell, inv_l_cell, n_cells, n_cells_tot, n_nei_cells)
        allocate(part_in_cell(0:n_cells_tot), cell_neigh_ls(n_cells_tot, n_nei_cells))
        call get_cell_neigh_list(n_cells_tot, n_cells, 3, n_nei_cells, cell_neigh_ls)
        call zigset(seed)
#ifdef _OPENMP
        nt = omp_get_max_threads()
        numth = nt
        allocate(seeds(0:nt - 1))
        do i = 0, nt - 1
            seeds(i) = seed + 7 * i + 1
        end do
        call par_zigset(nt, seeds, grainsize)
#else
        nt = 1
#endif
    end subroutine ref_setup

    subroutine ref_set_state(r0_, v_, a_, t_) bind(C, name="ref_set_state")
        real(c_double), intent(in) :: r0_(3, *), v_(3, *), a_(3, *)
        integer(c_int), intent(in) :: t_(*)
        integer :: i
        r0 = r0_(:, 1:n_part); v = v_(:, 1:n_part); a = a_(:, 1:n_part)
        a_type = t_(1:n_part)
        do i = 1, n_part                                  ! init_config.f90:326-329,374-378
            mass(i) = mass_type(a_type(i))
            inv_mass(i) = 1.0d0 / mass_type(a_type(i))
            ext_force(:, i) = 0.0d0
            if (a_type(i) /= 1) ext_force(:, i) = mass(i) * const_force(:)
        end do
        call neigh_list(r0, n_mon_tot, 3, n_cells, n_cells_tot, boundary, inv_l_cell, part_in_cell, r_nei, l_nei)
    end subroutine ref_set_state

    subroutine ref_get_state(r0_, v_, a_, f_) bind(C, name="ref_get_state")
        real(c_double), intent(out) :: r0_(3, *), v_(3, *), a_(3, *), f_(3, *)
        r0_(:, 1:n_part) = r0; v_(:, 1:n_part) = v; a_(:, 1:n_part) = a; f_(:, 1:n_part) = force
    end subroutine ref_get_state

    ! one routine on a zeroed force array; e returns the routine's energy
    subroutine ref_call(which, f_, e) bind(C, name="ref_call")
        integer(c_int), value :: which
        real(c_double), intent(out) :: f_(3, *), e
        force = 0.0d0
        e = 0.0d0
        select case (which)
        case (1)
            call fluid_fluid()
            e = v_fluid_fluid
        case (2)
            call fluid_wall(2)
            e = v_fluid_wall
        case (3)
            call intra_molec()
            e = v_intra_molec
        case (4)
            call bending(1)
            e = v_bend
        case (5)
            call constant_force(1)
        case (6)
            call check_fluid_fluid()
            e = v_fluid_fluid
        end select
        f_(:, 1:n_part) = force
    end subroutine ref_call

    ! md_main.f90:77-185 for n steps (ORIENTATION is not part of the path)
    subroutine ref_md_steps(n, t0, en) bind(C, name="ref_md_steps")
        integer(c_int), value :: n, t0
        real(c_double), intent(out) :: en(6)
        integer :: it
        do it = 1, n
            i_time = it
            r_time = dble(t0 + it - 1) * dt
            call verlet_positions()
            force(:, :) = 0.0d0
            call fluid_fluid()
            call fluid_wall(2)
            call wall_wall(2)
            call intra_molec()
            call bending(1)
            call constant_force(1)
            call verlet_velocities()
        end do
        en = (/ v_fluid_fluid, v_intra_molec, v_fluid_wall, v_bend, 0.0d0, t_fluid /)
    end subroutine ref_md_steps

    integer(c_int) function ref_num_threads() bind(C, name="ref_num_threads")
#ifdef _OPENMP
        ref_num_threads = omp_get_max_threads()
#else
        ref_num_threads = 1
#endif
    end function ref_num_threads

end module ref_harness
