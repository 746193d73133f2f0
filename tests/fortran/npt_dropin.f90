! Drop-in check of the Monte Carlo barostat and the reporters through the
! sucrose_gpu binding, from the reference's own language.  A Lennard-Jones
! system (108 atoms, rc 2.5) makes six mc_barostat moves with the
! reference's modules on the host (compute_potential: the pair loop over all
! pairs I < J) and six gpu_mc_barostat moves from the same state with the
! generator of random_number reset to the same seed; after each move the
! accept flag, the box and the positions must agree exactly.  Then the
! trajectory and checkpoint frames of the final state are written by the
! reference (report_trajectory, report_checkpoint) and by the device
! reporters into argument-named files, for the caller to compare.
program npt_dropin
    use iso_c_binding, only: c_ptr
    use iso_fortran_env, only: real64
    use system
    use lennard_jones
    use monte_carlo_barostat
    use trajectory_reporter
    use checkpoint_reporter
    use sucrose_gpu
    implicit none
    integer, parameter :: nmoves = 6
    real(real64), parameter :: rc = 2.5
    real(real64), allocatable :: rx0(:), ry0(:), rz0(:), rxr(:, :)
    real(real64) :: box0(3), boxr(3, nmoves)
    logical :: accr(nmoves), acc
    integer :: t, ns, k, nacc
    integer, allocatable :: seed(:)
    character(len=512) :: dir
    type(sx_step_spec) :: spec
    type(c_ptr) :: rep

    call get_command_argument(1, dir)
    call generate_fcc_lattice(3, 0.75_real64)
    call set_velocity_to_temperature(1.2_real64, seed=3)
    call set_periodic_boundary
    call initialize_system(.false.)
    call set_target_pressure(2.0_real64)
    temp = 1.2_real64
    rx0 = RX; ry0 = RY; rz0 = RZ; box0 = box
    allocate(rxr(N, nmoves))
    call random_seed(size=ns)
    allocate(seed(ns))
    do k = 1, ns
        seed(k) = 12345 + 7 * k
    enddo

    ! ---- the reference ----
    call random_seed(put=seed)
    nacc = 0
    do t = 1, nmoves
        call mc_barostat(host_potential, accr(t), 0.02_real64)
        rxr(:, t) = RX
        boxr(:, t) = box
        if (accr(t)) nacc = nacc + 1
    enddo
    write(*, '(A, I0, A, I0)') 'reference: accepted ', nacc, ' of ', nmoves

    ! ---- the binding, from the same state and seed ----
    RX = rx0; RY = ry0; RZ = rz0; box = box0
    call set_box(box0(1), box0(2), box0(3))
    temp = 1.2_real64
    call gpu_init(0)
    call gpu_push_system()
    call gpu_init_nlist(0.3_real64, rc)
    call gpu_set_target_pressure(2.0_real64)
    spec%nforces = 1
    spec%kinds = 0
    spec%kinds(1) = SX_LJ
    spec%params = 0.0
    spec%params(1:3, 1) = [1.0_real64, 1.0_real64, rc]
    spec%bonded = 0
    call random_seed(put=seed)
    do t = 1, nmoves
        call gpu_mc_barostat(spec, acc, 0.02_real64)
        call gpu_pull_system()
        if (acc .neqv. accr(t)) then
            write(*, '(A, I0)') 'accept differs at move ', t
            error stop 2
        endif
        if (any(box /= boxr(:, t)) .or. any(RX /= rxr(:, t))) then
            write(*, '(A, I0)') 'box or positions differ at move ', t
            error stop 3
        endif
    enddo
    write(*, '(A)') 'barostat moves agree'

    ! ---- frames of the final state: reference and device reporters ----
    call open_trajectory_file(31, trim(dir) // '/traj_ref.txt')
    call report_trajectory(31, 77)
    call close_trajectory_file(31)
    call open_checkpoint_file(32, trim(dir) // '/chk_ref.txt')
    call report_checkpoint(32)
    call close_checkpoint_file(32)
    rep = gpu_reporter_open(SX_REPORT_TRAJECTORY, trim(dir) // '/traj_gpu.txt')
    call gpu_report(rep, 77)
    call gpu_reporter_close(rep)
    rep = gpu_reporter_open(SX_REPORT_CHECKPOINT, trim(dir) // '/chk_gpu.txt')
    call gpu_report(rep, 0)
    call gpu_reporter_close(rep)
    call gpu_finalize()
    write(*, '(A)') 'NPT DROPIN OK'

contains

    ! compute_potential: initialize_forces, then the pair loop over all I < J
    real(real64) function host_potential()
        integer :: i, j
        real(real64) :: pe, f(3)
        call initialize_forces
        do i = 1, N - 1
            do j = i + 1, N
                call lj_calculate_forces(i, j, 1.0_real64, 1.0_real64, rc, pe, f)
                pot_eng = pot_eng + pe
                FX(i) = FX(i) + f(1); FY(i) = FY(i) + f(2); FZ(i) = FZ(i) + f(3)
                FX(j) = FX(j) - f(1); FY(j) = FY(j) - f(2); FZ(j) = FZ(j) - f(3)
            enddo
        enddo
        host_potential = pot_eng
    end function host_potential

end program npt_dropin
