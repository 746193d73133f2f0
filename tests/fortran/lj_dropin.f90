! Drop-in check from the reference's own language: the C1 Lennard-Jones NVE
! run (864 atoms on an FCC lattice at density 0.8442, T* = 0.722 from seed 15,
! rc = 2.5, skin 0.3, dt = 0.005) driven twice from the same initial state --
! once by the reference's modules on the host (README.md:83-115 with the
! reference's Verlet list), once through the sucrose_gpu binding on the GPU --
! and the per-step potential and kinetic energies compared.
! Prints one line per checkpoint and a final verdict; exit status 0 on success.
program lj_dropin
    use iso_fortran_env, only: real64
    use system
    use neighbour_list
    use verlet
    use lennard_jones
    use sucrose_gpu
    implicit none
    integer, parameter :: nsteps = 1000
    real(real64), parameter :: dt = 0.005, rc = 2.5, skin = 0.3
    real(real64) :: pe_ref(0:nsteps), ke_ref(0:nsteps), pe_gpu(0:nsteps), ke_gpu(0:nsteps)
    real(real64) :: lj(3), d100, d1000, e0r, e0g, driftr, driftg
    real(real64), allocatable :: rx_ref(:)
    integer :: t

    lj = [1.0_real64, 1.0_real64, rc]
    call generate_fcc_lattice(6, 0.8442_real64)
    call set_velocity_to_temperature(0.722_real64, seed=15)
    call set_periodic_boundary
    call initialize_system(.false.)

    ! device copy of the identical initial state
    call gpu_init(0)
    call gpu_push_system()

    ! ---- the reference on the host ----
    call init_nlist(skin, rc)
    call host_forces()
    call calculate_thermo
    pe_ref(0) = pot_eng; ke_ref(0) = kin_eng
    do t = 1, nsteps
        call vv_initial_step(dt)
        if (needs_rebuild()) call build_nlist
        call host_forces()
        call vv_final_step(dt)
        call calculate_thermo
        pe_ref(t) = pot_eng; ke_ref(t) = kin_eng
    enddo
    rx_ref = RX

    ! ---- the same loop on the GPU through the binding ----
    call gpu_init_nlist(skin, rc)
    call gpu_initialize_forces()
    call gpu_calculate_forces(SX_LJ, lj)
    call gpu_calculate_thermo()
    call gpu_pull_system()
    pe_gpu(0) = pot_eng; ke_gpu(0) = kin_eng
    do t = 1, nsteps
        call gpu_vv_initial_step(dt)
        call gpu_update_nlist()
        call gpu_initialize_forces()
        call gpu_calculate_forces(SX_LJ, lj)
        call gpu_vv_final_step(dt)
        call gpu_calculate_thermo()
        call gpu_pull_system()
        pe_gpu(t) = pot_eng; ke_gpu(t) = kin_eng
    enddo
    call gpu_finalize()

    d100 = max(maxval(abs(pe_gpu(0:100) - pe_ref(0:100))), maxval(abs(ke_gpu(0:100) - ke_ref(0:100)))) &
           / abs(pe_ref(0))
    d1000 = max(maxval(abs(pe_gpu - pe_ref)), maxval(abs(ke_gpu - ke_ref))) / abs(pe_ref(0))
    e0r = pe_ref(0) + ke_ref(0); e0g = pe_gpu(0) + ke_gpu(0)
    driftr = maxval(abs(pe_ref + ke_ref - e0r)) / abs(e0r)
    driftg = maxval(abs(pe_gpu + ke_gpu - e0g)) / abs(e0g)
    write(*, '(A, I0, A, ES12.5, A, ES12.5)') 'N=', N, ' E0(ref)=', e0r, ' E0(gpu)=', e0g
    write(*, '(A, ES10.3)') 'max |dE|/|PE0| over steps 0-100:  ', d100
    write(*, '(A, ES10.3)') 'max |dE|/|PE0| over steps 0-1000: ', d1000
    write(*, '(A, ES10.3, A, ES10.3)') 'NVE drift max|E-E0|/|E0|: reference ', driftr, '  gpu ', driftg
    write(*, '(A, ES10.3)') 'final positions max |dx|: ', maxval(abs(RX - rx_ref))
    if (d100 < 1e-10_real64 .and. d1000 < 1e-5_real64 .and. abs(driftg - driftr) < 1e-4_real64) then
        write(*, '(A)') 'DROPIN OK'
    else
        write(*, '(A)') 'DROPIN FAIL'
        error stop 1
    endif

contains

    ! the README pair loop over the reference's neighbour list
    subroutine host_forces()
        integer :: I, K, J
        real(real64) :: pe, f(3)
        call initialize_forces
        do I = 1, N - 1
            do K = offset(I), offset(I + 1) - 1
                J = nlist(K)
                call lj_calculate_forces(I, J, lj(1), lj(2), lj(3), pe, f)
                pot_eng = pot_eng + pe
                FX(I) = FX(I) + f(1); FY(I) = FY(I) + f(2); FZ(I) = FZ(I) + f(3)
                FX(J) = FX(J) - f(1); FY(J) = FY(J) - f(2); FZ(J) = FZ(J) - f(3)
            enddo
        enddo
    end subroutine host_forces

end program lj_dropin
