! Test harness around the SUCROSE reference modules -- TEST INFRASTRUCTURE.
! Compiled together with the reference's own fortran/*.f90 (read where they lie
! under the reference checkout) into oracle/_ref/libsucrose_ref.so by
! oracle/Makefile.  It only moves data in and out of the reference's module
! globals and calls the reference routines unchanged, so that the C oracle can
! be checked against the reference on identical inputs.
module ref_harness
    use iso_c_binding
    use philox_rng
    use xmath
    use system
    use neighbour_list
    use constraints
    use lennard_jones
    use wca
    use morse
    use gay_berne
    use ecp
    use bonded
    use debye_huckel
    use coulumbic
    use verlet
    use quaternion
    use nose_hoover
    use langevin
    use monte_carlo_barostat
    use dipole
    use reporters
    implicit none

    integer, parameter :: NFIELD = 37
    ! Lennard-Jones parameters of harness_potential (mc_barostat's compute_potential)
    real(real64) :: hp_sigma = 1.0, hp_eps = 1.0, hp_rc = 2.5

contains

    subroutine free_if(a)
        real(real64), allocatable, intent(inout) :: a(:)
        if (allocated(a)) deallocate(a)
    end subroutine free_if

    ! Allocate the module arrays for n particles (set_num_atoms, system.f90:43).
    subroutine ref_alloc(n, nrigid) bind(C, name="ref_alloc")
        integer(c_int), value :: n, nrigid
        call free_if(mass); call free_if(inv_mass); call free_if(charge)
        call free_if(sig_x); call free_if(sig_y); call free_if(sig_z)
        call free_if(eps_x); call free_if(eps_y); call free_if(eps_z)
        call free_if(Ixx); call free_if(Iyy); call free_if(Izz)
        call free_if(Ixx_inv); call free_if(Iyy_inv); call free_if(Izz_inv)
        call free_if(RX); call free_if(RY); call free_if(RZ)
        call free_if(VX); call free_if(VY); call free_if(VZ)
        call free_if(FX); call free_if(FY); call free_if(FZ)
        call free_if(LX); call free_if(LY); call free_if(LZ)
        call free_if(TX); call free_if(TY); call free_if(TZ)
        call free_if(QW); call free_if(QX); call free_if(QY); call free_if(QZ)
        call free_if(RX_old); call free_if(RY_old); call free_if(RZ_old)
        if (allocated(rtype)) deallocate(rtype)
        if (allocated(BBI)) deallocate(BBI)
        if (allocated(BBJ)) deallocate(BBJ)
        if (allocated(BB_len)) deallocate(BB_len)
        if (nrigid > 1) then
            call set_num_atoms(n / nrigid, nrigid)
        else
            call set_num_atoms(n)
        endif
        allocate(RX_old(N), RY_old(N), RZ_old(N))
        RX_old = 0.0; RY_old = 0.0; RZ_old = 0.0
        NC = 0
        allocate(BBI(1), BBJ(1), BB_len(1))
    end subroutine ref_alloc

    subroutine ref_set_box(box3, per) bind(C, name="ref_set_box")
        real(c_double), intent(in) :: box3(3)
        integer(c_int), value :: per
        box = box3
        periodic = (per /= 0)
        density = dble(N / Nrigid) / product(box)
    end subroutine ref_set_box

    ! Field order shared with oracle/oracle.py (FIELDS).
    subroutine ref_set_state(n_in, buf) bind(C, name="ref_set_state")
        integer(c_int), value :: n_in
        real(c_double), intent(in) :: buf(n_in, NFIELD)
        RX = buf(:, 1); RY = buf(:, 2); RZ = buf(:, 3)
        VX = buf(:, 4); VY = buf(:, 5); VZ = buf(:, 6)
        FX = buf(:, 7); FY = buf(:, 8); FZ = buf(:, 9)
        QW = buf(:, 10); QX = buf(:, 11); QY = buf(:, 12); QZ = buf(:, 13)
        LX = buf(:, 14); LY = buf(:, 15); LZ = buf(:, 16)
        TX = buf(:, 17); TY = buf(:, 18); TZ = buf(:, 19)
        mass = buf(:, 20); inv_mass = buf(:, 21); charge = buf(:, 22)
        sig_x = buf(:, 23); sig_y = buf(:, 24); sig_z = buf(:, 25)
        eps_x = buf(:, 26); eps_y = buf(:, 27); eps_z = buf(:, 28)
        Ixx = buf(:, 29); Iyy = buf(:, 30); Izz = buf(:, 31)
        Ixx_inv = buf(:, 32); Iyy_inv = buf(:, 33); Izz_inv = buf(:, 34)
        RX_old = buf(:, 35); RY_old = buf(:, 36); RZ_old = buf(:, 37)
    end subroutine ref_set_state

    subroutine ref_get_state(n_in, buf) bind(C, name="ref_get_state")
        integer(c_int), value :: n_in
        real(c_double), intent(out) :: buf(n_in, NFIELD)
        buf(:, 1) = RX; buf(:, 2) = RY; buf(:, 3) = RZ
        buf(:, 4) = VX; buf(:, 5) = VY; buf(:, 6) = VZ
        buf(:, 7) = FX; buf(:, 8) = FY; buf(:, 9) = FZ
        buf(:, 10) = QW; buf(:, 11) = QX; buf(:, 12) = QY; buf(:, 13) = QZ
        buf(:, 14) = LX; buf(:, 15) = LY; buf(:, 16) = LZ
        buf(:, 17) = TX; buf(:, 18) = TY; buf(:, 19) = TZ
        buf(:, 20) = mass; buf(:, 21) = inv_mass; buf(:, 22) = charge
        buf(:, 23) = sig_x; buf(:, 24) = sig_y; buf(:, 25) = sig_z
        buf(:, 26) = eps_x; buf(:, 27) = eps_y; buf(:, 28) = eps_z
        buf(:, 29) = Ixx; buf(:, 30) = Iyy; buf(:, 31) = Izz
        buf(:, 32) = Ixx_inv; buf(:, 33) = Iyy_inv; buf(:, 34) = Izz_inv
        buf(:, 35) = RX_old; buf(:, 36) = RY_old; buf(:, 37) = RZ_old
    end subroutine ref_get_state

    ! Scalars: pot_eng kin_eng kin_rot temp pressure density virial virial_bonded
    !          target_temp dof HB(3) eta(3) p_eta(3)   (19 values)
    subroutine ref_get_scalars(s) bind(C, name="ref_get_scalars")
        real(c_double), intent(out) :: s(19)
        s(1) = pot_eng; s(2) = kin_eng; s(3) = kin_rot; s(4) = temp
        s(5) = pressure; s(6) = density; s(7) = virial; s(8) = virial_bonded
        s(9) = target_temp; s(10) = dble(dof)
        s(11:13) = HB; s(14:16) = eta; s(17:19) = p_eta
    end subroutine ref_get_scalars

    subroutine ref_set_scalars(s) bind(C, name="ref_set_scalars")
        real(c_double), intent(in) :: s(19)
        pot_eng = s(1); kin_eng = s(2); kin_rot = s(3); temp = s(4)
        pressure = s(5); density = s(6); virial = s(7); virial_bonded = s(8)
        target_temp = s(9); dof = nint(s(10))
        HB = s(11:13); eta = s(14:16); p_eta = s(17:19)
        target_temp_set = .TRUE.
    end subroutine ref_set_scalars

    subroutine ref_set_bonds(nc_in, bi, bj, blen) bind(C, name="ref_set_bonds")
        integer(c_int), value :: nc_in
        integer(c_int), intent(in) :: bi(nc_in), bj(nc_in)
        real(c_double), intent(in) :: blen(nc_in)
        if (allocated(BBI)) deallocate(BBI)
        if (allocated(BBJ)) deallocate(BBJ)
        if (allocated(BB_len)) deallocate(BB_len)
        NC = nc_in
        allocate(BBI(NC), BBJ(NC), BB_len(NC))
        BBI = bi; BBJ = bj; BB_len = blen
    end subroutine ref_set_bonds

    subroutine ref_initialize_system(remove_com) bind(C, name="ref_initialize_system")
        integer(c_int), value :: remove_com
        remove_com_flag = (remove_com /= 0)
        call initialize_system(.FALSE.)
    end subroutine ref_initialize_system

    subroutine ref_initialize_forces() bind(C, name="ref_initialize_forces")
        call initialize_forces
    end subroutine ref_initialize_forces

    subroutine ref_calculate_thermo() bind(C, name="ref_calculate_thermo")
        call calculate_thermo
    end subroutine ref_calculate_thermo

    subroutine ref_set_velocity_to_temperature(t, seed, remove_com) bind(C, name="ref_set_velocity_to_temperature")
        real(c_double), value :: t
        integer(c_int), value :: seed, remove_com
        call set_velocity_to_temperature(t, seed, remove_com /= 0)
    end subroutine ref_set_velocity_to_temperature

    subroutine ref_generate_fcc_lattice(uc, dens) bind(C, name="ref_generate_fcc_lattice")
        integer(c_int), value :: uc
        real(c_double), value :: dens
        call generate_fcc_lattice(uc, dens)
    end subroutine ref_generate_fcc_lattice

    ! ---- Philox ----
    subroutine ref_philox_uniform(ctr, out) bind(C, name="ref_philox_uniform")
        integer(c_int32_t), intent(in) :: ctr(4)
        real(c_double), intent(out) :: out(4)
        call philox_uniform(ctr, out)
    end subroutine ref_philox_uniform

    subroutine ref_philox_normal(c1, c2, c3, c4, out) bind(C, name="ref_philox_normal")
        integer(c_int32_t), value :: c1, c2, c3, c4
        real(c_double), intent(out) :: out(3)
        call philox_normal_3seq(c1, c2, c3, c4, out(1), out(2), out(3))
    end subroutine ref_philox_normal

    ! ---- 3x3 helpers ----
    subroutine ref_cholesky_solve(a, b, out) bind(C, name="ref_cholesky_solve")
        real(c_double), intent(in) :: a(3, 3), b(3)
        real(c_double), intent(out) :: out(3)
        out = cholesky_solve(a, b)
    end subroutine ref_cholesky_solve

    function ref_brent(ar, br, r) result(lam) bind(C, name="ref_brent")
        real(c_double), intent(in) :: ar(3, 3), br(3, 3), r(3)
        real(c_double) :: lam
        type(ecp_pair) :: pr
        pr % AR = ar; pr % BR = br; pr % RIJ = r
        pr % AE = ar; pr % BE = br
        lam = brent(optim_dist, pr)
    end function ref_brent

    ! ---- pair kernels: kind codes as in include/sucrose.h ----
    subroutine ref_pair(kind, i1, j1, p, pe, f, ti, tj) bind(C, name="ref_pair")
        integer(c_int), value :: kind, i1, j1
        real(c_double), intent(in) :: p(*)
        real(c_double), intent(out) :: pe, f(3), ti(3), tj(3)
        pe = 0.0; f = 0.0; ti = 0.0; tj = 0.0
        select case (kind)
        case (1)
            call lj_calculate_forces(i1, j1, p(1), p(2), p(3), pe, f)
        case (2)
            call lj12_calculate_forces(i1, j1, p(1), p(2), p(3), pe, f)
        case (3)
            call ljw_calculate_forces(i1, j1, p(1), p(2), p(3), pe, f, p(4))
        case (4)
            call wca_calculate_forces(i1, j1, p(1), p(2), pe, f)
        case (5)
            call morse_calculate_forces(i1, j1, p(1), p(2), p(3), p(4), pe, f)
        case (6)
            call gb_calculate_forces(i1, j1, p(1), pe, f, ti, tj)
        case (7)
            call gbr_calculate_forces(i1, j1, pe, f, ti, tj)
        case (8)
            call ecp_calculate_forces(i1, j1, p(1), pe, f, ti, tj)
        case (9)
            call ecpr_calculate_forces(i1, j1, pe, f, ti, tj)
        case (10)
            call dh_calculate_forces(i1, j1, p(1), p(2), p(3), pe, f)
        case (11)   ! p = cutoff, (mu, nu fixed by the reference), wf
            call gbw_calculate_forces(i1, j1, p(1), pe, f, ti, tj, p(4))
        case (12)   ! p = cutoff, -, -, chirality
            call gbc_calculate_forces(i1, j1, p(1), pe, f, ti, tj, p(4))
        case (13)   ! p = cutoff, -, -, wf, chirality
            call gbwc_calculate_forces(i1, j1, p(1), pe, f, ti, tj, p(4), p(5))
        case (14)   ! p = cutoff, wf
            call ecpw_calculate_forces(i1, j1, p(1), pe, f, ti, tj, p(2))
        case (15)   ! p = cutoff, chirality
            call ecpc_calculate_forces(i1, j1, p(1), pe, f, ti, tj, p(2))
        case (16)   ! p = cutoff, wf, chirality
            call ecpwc_calculate_forces(i1, j1, p(1), pe, f, ti, tj, p(2), p(3))
        case (18)   ! p = cutoff, dielectric
            call coulombic_calculate_forces(i1, j1, p(1), p(2), pe, f)
        end select
    end subroutine ref_pair

    ! ---- bonded: kind 1 bend, 2 cosine_bend, 3 dih, 4 cosine_dih ----
    subroutine ref_bonded(kind, i1, j1, k1, l1, coeff, angle, pe, f) bind(C, name="ref_bonded")
        integer(c_int), value :: kind, i1, j1, k1, l1
        real(c_double), value :: coeff, angle
        real(c_double), intent(out) :: pe, f(3, 4)
        f = 0.0; pe = 0.0
        select case (kind)
        case (1)
            call bend_calculate_forces(i1, j1, k1, pe, f(:, 1), f(:, 2), f(:, 3), coeff)
        case (2)
            call cosine_bend_calculate_forces(i1, j1, k1, pe, f(:, 1), f(:, 2), f(:, 3), coeff, angle)
        case (3)
            call dih_calculate_forces(i1, j1, k1, l1, pe, f(:, 1), f(:, 2), f(:, 3), f(:, 4), coeff)
        case (4)
            call cosine_dih_calculate_forces(i1, j1, k1, l1, pe, f(:, 1), f(:, 2), f(:, 3), f(:, 4), coeff, angle)
        case (5)
            call angle_bend_calculate_forces(i1, j1, k1, pe, f(:, 1), f(:, 2), f(:, 3), coeff, angle)
        case (6)
            call angle_dih_calculate_forces(i1, j1, k1, l1, pe, f(:, 1), f(:, 2), f(:, 3), f(:, 4), coeff, angle)
        end select
    end subroutine ref_bonded

    ! ---- integrators: kind 1 vv, 2 vvc, 3 qq, 4 nht, 5 nhtc, 6 lgv, 7 lgvc;
    !      phase 0 initial, 1 final ----
    subroutine ref_integrate(kind, phase, dt, step, damping, has_coupling, coupling) &
            bind(C, name="ref_integrate")
        integer(c_int), value :: kind, phase, step, has_coupling
        real(c_double), value :: dt, damping, coupling
        select case (kind)
        case (1)
            if (phase == 0) then
                call vv_initial_step(dt)
            else
                call vv_final_step(dt)
            endif
        case (2)
            if (phase == 0) then
                call vvc_initial_step(dt)
            else
                call vvc_final_step(dt)
            endif
        case (3)
            if (phase == 0) then
                call qq_initial_step(dt)
            else
                call qq_final_step(dt)
            endif
        case (4)
            if (phase == 0) then
                if (has_coupling /= 0) then
                    call nht_initial_step(dt, coupling)
                else
                    call nht_initial_step(dt)
                endif
            else
                if (has_coupling /= 0) then
                    call nht_final_step(dt, coupling)
                else
                    call nht_final_step(dt)
                endif
            endif
        case (5)
            if (phase == 0) then
                call nhtc_initial_step(dt)
            else
                call nhtc_final_step(dt)
            endif
        case (6)
            if (phase == 0) then
                call lgv_initial_step(dt, step, damping)
            else
                call lgv_final_step(dt, damping)
            endif
        case (7)
            if (phase == 0) then
                call lgvc_initial_step(dt, step, damping)
            else
                call lgvc_final_step(dt, damping)
            endif
        end select
    end subroutine ref_integrate

    subroutine ref_constraints(phase, dt) bind(C, name="ref_constraints")
        integer(c_int), value :: phase
        real(c_double), value :: dt
        if (phase == 0) then
            call apply_constraints_a(dt)
        else
            call apply_constraints_b(dt)
        endif
    end subroutine ref_constraints

    ! ---- neighbour list (one init_nlist per process: the module keeps
    !      private arrays that only init_nlist allocates) ----
    subroutine ref_nlist_init(skin, cutoff) bind(C, name="ref_nlist_init")
        real(c_double), value :: skin, cutoff
        call init_nlist(skin, cutoff)
    end subroutine ref_nlist_init

    subroutine ref_nlist_build() bind(C, name="ref_nlist_build")
        call build_nlist
    end subroutine ref_nlist_build

    function ref_needs_rebuild() result(r) bind(C, name="ref_needs_rebuild")
        integer(c_int) :: r
        r = 0
        if (needs_rebuild()) r = 1
    end function ref_needs_rebuild

    ! offset_out(1:N) as in the reference (1-based, offset(N) = K+1); nnz = K
    subroutine ref_nlist_get(cap, offset_out, nbr_out, nnz) bind(C, name="ref_nlist_get")
        integer(c_long), value :: cap
        integer(c_int), intent(out) :: offset_out(N)
        integer(c_int), intent(out) :: nbr_out(cap)
        integer(c_long), intent(out) :: nnz
        integer :: k
        offset_out = offset
        nnz = offset(N) - 1
        do k = 1, int(min(int(nnz, c_long), cap))
            nbr_out(k) = nlist(k)
        enddo
    end subroutine ref_nlist_get

    ! ---- the README driver loop (README.md:83-115) for the workloads the
    !      benchmark times.  integ: 1 vv, 4 nht (+qq), 6 lgv.  force kinds as
    !      in ref_pair.  The neighbour list must have been initialised. ----
    subroutine ref_run_steps(integ, fkind, p, nsteps, step0, dt, damping) bind(C, name="ref_run_steps")
        integer(c_int), value :: integ, fkind, nsteps, step0
        real(c_double), intent(in) :: p(*)
        real(c_double), value :: dt, damping
        integer :: t, I, K, J
        real(real64) :: pe, f(3), ti(3), tj(3)
        do t = 1, nsteps
            select case (integ)
            case (1)
                call vv_initial_step(dt)
            case (4)
                call nht_initial_step(dt)
                call qq_initial_step(dt)
            case (6)
                call lgv_initial_step(dt, step0 + t - 1, damping)
            end select
            if (needs_rebuild()) call build_nlist
            call initialize_forces
            do I = 1, N - 1
                do K = offset(I), offset(I + 1) - 1
                    J = nlist(K)
                    call ref_pair(fkind, I, J, p, pe, f, ti, tj)
                    pot_eng = pot_eng + pe
                    FX(I) = FX(I) + f(1); FY(I) = FY(I) + f(2); FZ(I) = FZ(I) + f(3)
                    FX(J) = FX(J) - f(1); FY(J) = FY(J) - f(2); FZ(J) = FZ(J) - f(3)
                    TX(I) = TX(I) + ti(1); TY(I) = TY(I) + ti(2); TZ(I) = TZ(I) + ti(3)
                    TX(J) = TX(J) + tj(1); TY(J) = TY(J) + tj(2); TZ(J) = TZ(J) + tj(3)
                enddo
            enddo
            select case (integ)
            case (1)
                call vv_final_step(dt)
            case (4)
                call nht_final_step(dt)
                call qq_final_step(dt)
            case (6)
                call lgv_final_step(dt, damping)
            end select
            call calculate_thermo
        enddo
    end subroutine ref_run_steps

    ! ---- the charged-chain loop: lgvc_initial_step (with apply_constraints_a),
    !      the rebuild check, one pass over the list calling
    !      wca_calculate_forces and dh_calculate_forces for each pair, the
    !      cosine_bend_ and dih_ terms, lgvc_final_step (with
    !      apply_constraints_b), calculate_thermo.  pw = (sigma, epsilon),
    !      pd = (cutoff, db_len, dielectric); bend(3, nbend) and dih(4, ndih)
    !      hold 1-based atoms. ----
    subroutine ref_run_chain_steps(nsteps, step0, dt, damping, pw, pd, nbend, bend, kbend, abend, ndih, dih, kdih) &
            bind(C, name="ref_run_chain_steps")
        integer(c_int), value :: nsteps, step0, nbend, ndih
        real(c_double), value :: dt, damping, kbend, abend, kdih
        real(c_double), intent(in) :: pw(2), pd(3)
        integer(c_int), intent(in) :: bend(3, nbend), dih(4, ndih)
        integer :: t, I, K, J, m
        real(real64) :: pe, f(3), f4(3, 4)
        do t = 1, nsteps
            call lgvc_initial_step(dt, step0 + t - 1, damping)
            if (needs_rebuild()) call build_nlist
            call initialize_forces
            do I = 1, N - 1
                do K = offset(I), offset(I + 1) - 1
                    J = nlist(K)
                    call wca_calculate_forces(I, J, pw(1), pw(2), pe, f)
                    pot_eng = pot_eng + pe
                    FX(I) = FX(I) + f(1); FY(I) = FY(I) + f(2); FZ(I) = FZ(I) + f(3)
                    FX(J) = FX(J) - f(1); FY(J) = FY(J) - f(2); FZ(J) = FZ(J) - f(3)
                    call dh_calculate_forces(I, J, pd(1), pd(2), pd(3), pe, f)
                    pot_eng = pot_eng + pe
                    FX(I) = FX(I) + f(1); FY(I) = FY(I) + f(2); FZ(I) = FZ(I) + f(3)
                    FX(J) = FX(J) - f(1); FY(J) = FY(J) - f(2); FZ(J) = FZ(J) - f(3)
                enddo
            enddo
            do m = 1, nbend
                f4 = 0.0
                call cosine_bend_calculate_forces(bend(1, m), bend(2, m), bend(3, m), pe, f4(:, 1), f4(:, 2), &
                                                  f4(:, 3), kbend, abend)
                pot_eng = pot_eng + pe
                do K = 1, 3
                    FX(bend(K, m)) = FX(bend(K, m)) + f4(1, K)
                    FY(bend(K, m)) = FY(bend(K, m)) + f4(2, K)
                    FZ(bend(K, m)) = FZ(bend(K, m)) + f4(3, K)
                enddo
            enddo
            do m = 1, ndih
                f4 = 0.0
                call dih_calculate_forces(dih(1, m), dih(2, m), dih(3, m), dih(4, m), pe, f4(:, 1), f4(:, 2), &
                                          f4(:, 3), f4(:, 4), kdih)
                pot_eng = pot_eng + pe
                do K = 1, 4
                    FX(dih(K, m)) = FX(dih(K, m)) + f4(1, K)
                    FY(dih(K, m)) = FY(dih(K, m)) + f4(2, K)
                    FZ(dih(K, m)) = FZ(dih(K, m)) + f4(3, K)
                enddo
            enddo
            call lgvc_final_step(dt, damping)
            call calculate_thermo
        enddo
    end subroutine ref_run_chain_steps

    ! ---- reporters (reporters.f90): kind 1 state data, 2 trajectory,
    !      3 config, 4 checkpoint; one frame written to a new file ----
    function cstr(c, n) result(f)
        character(kind=c_char), intent(in) :: c(*)
        integer, intent(in) :: n
        character(len=n) :: f
        integer :: k
        do k = 1, n
            f(k:k) = c(k)
        enddo
    end function cstr

    subroutine ref_report(kind, fname, flen, step, rt) bind(C, name="ref_report")
        integer(c_int), value :: kind, flen, step
        character(kind=c_char), intent(in) :: fname(*)
        integer(c_int), intent(in) :: rt(*)
        integer, parameter :: u = 91
        rtype = rt(1:N)
        select case (kind)
        case (1)
            call open_state_data_file(u, cstr(fname, flen))
            call report_state_data(u, step)
            call close_state_data_file(u)
        case (2)
            call open_trajectory_file(u, cstr(fname, flen))
            call report_trajectory(u, step)
            call close_trajectory_file(u)
        case (3)
            call open_confwriter_file(u, cstr(fname, flen))
            call report_confwriter(u)
            call close_confwriter_file(u)
        case (4)
            call open_checkpoint_file(u, cstr(fname, flen))
            call report_checkpoint(u)
            call close_checkpoint_file(u)
        end select
    end subroutine ref_report

    ! load_checkpoint (reporters.f90:186-225) into freshly allocated module
    ! arrays; returns N and NC
    subroutine ref_load_checkpoint(fname, flen, n_out, nc_out, cfg) bind(C, name="ref_load_checkpoint")
        integer(c_int), value :: flen
        character(kind=c_char), intent(in) :: fname(*)
        integer(c_int), intent(out) :: n_out, nc_out, cfg
        integer, parameter :: u = 92
        call free_if(mass); call free_if(inv_mass); call free_if(charge)
        call free_if(sig_x); call free_if(sig_y); call free_if(sig_z)
        call free_if(eps_x); call free_if(eps_y); call free_if(eps_z)
        call free_if(Ixx); call free_if(Iyy); call free_if(Izz)
        call free_if(Ixx_inv); call free_if(Iyy_inv); call free_if(Izz_inv)
        call free_if(RX); call free_if(RY); call free_if(RZ)
        call free_if(VX); call free_if(VY); call free_if(VZ)
        call free_if(FX); call free_if(FY); call free_if(FZ)
        call free_if(LX); call free_if(LY); call free_if(LZ)
        call free_if(TX); call free_if(TY); call free_if(TZ)
        call free_if(QW); call free_if(QX); call free_if(QY); call free_if(QZ)
        call free_if(RX_old); call free_if(RY_old); call free_if(RZ_old)
        if (allocated(rtype)) deallocate(rtype)
        if (allocated(BBI)) deallocate(BBI)
        if (allocated(BBJ)) deallocate(BBJ)
        if (allocated(BB_len)) deallocate(BB_len)
        open(unit=u, file=cstr(fname, flen), action='read')
        call load_checkpoint(u)
        close(u)
        n_out = N
        nc_out = NC
        cfg = config_index
    end subroutine ref_load_checkpoint

    subroutine ref_get_types(rt) bind(C, name="ref_get_types")
        integer(c_int), intent(out) :: rt(*)
        rt(1:N) = rtype
    end subroutine ref_get_types

    subroutine ref_get_bonds(bi, bj, blen) bind(C, name="ref_get_bonds")
        integer(c_int), intent(out) :: bi(*), bj(*)
        real(c_double), intent(out) :: blen(*)
        if (NC > 0) then
            bi(1:NC) = BBI(1:NC); bj(1:NC) = BBJ(1:NC); blen(1:NC) = BB_len(1:NC)
        endif
    end subroutine ref_get_bonds

    subroutine ref_get_box(b) bind(C, name="ref_get_box")
        real(c_double), intent(out) :: b(3)
        b = box
    end subroutine ref_get_box

    ! compute_potential for mc_barostat: initialize_forces, then the
    ! Lennard-Jones pair loop over all pairs I < J; returns pot_eng
    real(real64) function harness_potential()
        integer :: I, J
        real(real64) :: pe, fij(3)
        call initialize_forces
        do I = 1, N - 1
            do J = I + 1, N
                call lj_calculate_forces(I, J, hp_sigma, hp_eps, hp_rc, pe, fij)
                pot_eng = pot_eng + pe
                FX(I) = FX(I) + fij(1); FY(I) = FY(I) + fij(2); FZ(I) = FZ(I) + fij(3)
                FX(J) = FX(J) - fij(1); FY(J) = FY(J) - fij(2); FZ(J) = FZ(J) - fij(3)
            enddo
        enddo
        harness_potential = pot_eng
    end function harness_potential

    ! mc_barostat (integrators.f90:864-945) with harness_potential; the
    ! generator of random_number is seeded from `seed` and the two numbers
    ! the move draws are returned in u (drawn ahead from the same seed)
    subroutine ref_mc_barostat(seed, pres, volume_change, sigma, eps, rc, u, acc) bind(C, name="ref_mc_barostat")
        integer(c_int), value :: seed
        real(c_double), value :: pres, volume_change, sigma, eps, rc
        real(c_double), intent(out) :: u(2)
        integer(c_int), intent(out) :: acc
        integer :: ns, k
        integer, allocatable :: sd(:)
        logical :: accept
        hp_sigma = sigma; hp_eps = eps; hp_rc = rc
        call random_seed(size=ns)
        allocate(sd(ns))
        do k = 1, ns
            sd(k) = seed + 7919 * k
        enddo
        call random_seed(put=sd)
        call random_number(u(1))
        call random_number(u(2))
        call random_seed(put=sd)
        call set_target_pressure(pres)
        accept = .false.
        call mc_barostat(harness_potential, accept, volume_change)
        acc = merge(1, 0, accept)
    end subroutine ref_mc_barostat

    ! ---- module dipole (force.f90:2669-2767) ----
    subroutine ref_init_reaction_field() bind(C, name="ref_init_reaction_field")
        if (allocated(DIPX)) deallocate(DIPX, DIPY, DIPZ, RFLX, RFLY, RFLZ)
        call init_reaction_field
    end subroutine ref_init_reaction_field

    subroutine ref_dipole_pair(i1, j1, cutoff, dielectric, pe, f, ti, tj) bind(C, name="ref_dipole_pair")
        integer(c_int), value :: i1, j1
        real(c_double), value :: cutoff, dielectric
        real(c_double), intent(out) :: pe, f(3), ti(3), tj(3)
        call dipole_calculate_forces(i1, j1, cutoff, dielectric, pe, f, ti, tj)
    end subroutine ref_dipole_pair

    subroutine ref_eval_reaction_field() bind(C, name="ref_eval_reaction_field")
        call eval_reaction_field
    end subroutine ref_eval_reaction_field

    subroutine ref_get_dipole(dip, rfl) bind(C, name="ref_get_dipole")
        real(c_double), intent(out) :: dip(N, 3), rfl(N, 3)
        dip(:, 1) = DIPX; dip(:, 2) = DIPY; dip(:, 3) = DIPZ
        rfl(:, 1) = RFLX; rfl(:, 2) = RFLY; rfl(:, 3) = RFLZ
    end subroutine ref_get_dipole

    ! ---- composite rigid bodies (system.f90:743-881) ----
    subroutine ref_copy_transform(from1, to1, xo, yo, zo, set_inertia) bind(C, name="ref_copy_transform")
        integer(c_int), value :: from1, to1, set_inertia
        real(c_double), value :: xo, yo, zo
        if (set_inertia < 0) then
            call copy_transform(from1, to1, xo, yo, zo)
        else
            call copy_transform(from1, to1, xo, yo, zo, set_inertia /= 0)
        endif
    end subroutine ref_copy_transform

    subroutine ref_transfer_forces(from1, to1) bind(C, name="ref_transfer_forces")
        integer(c_int), value :: from1, to1
        call transfer_forces(from1, to1)
    end subroutine ref_transfer_forces

end module ref_harness
