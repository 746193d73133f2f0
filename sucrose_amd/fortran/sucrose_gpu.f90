! sucrose_gpu -- Fortran binding of libsucrose_hip (include/sucrose.h) for the
! SUCROSE Fortran modules.
!
! The interfaces are bind(C) declarations of the C ABI.  The gpu_* wrappers
! keep the reference routines' names and arguments (vv_initial_step(dt) ->
! gpu_vv_initial_step(dt), ...) and stop the program on a non-zero status, as
! the reference does on its own error paths (with a non-zero exit code).  State moves between the reference's
! module arrays (module system, fortran/system.f90:7-33) and the device only
! in gpu_push_system / gpu_pull_system; between those calls the whole loop of
! README.md:83-115 runs on the GPU.
module sucrose_gpu
    use iso_c_binding
    use iso_fortran_env, only: real64
    use system
    implicit none
    private

    type(c_ptr), save :: ctx = c_null_ptr

    ! pair kinds (sx_kind) and integrators (sx_integrator)
    integer(c_int), parameter, public :: SX_LJ = 1, SX_LJ12 = 2, SX_LJW = 3, SX_WCA = 4, SX_MORSE = 5, &
        SX_GB = 6, SX_GBR = 7, SX_ECP = 8, SX_ECPR = 9, SX_DH = 10
    integer(c_int), parameter, public :: SX_INT_VV = 1, SX_INT_NHT = 4, SX_INT_LGV = 6, SX_INT_LGVC = 7
    ! scalar ids (sx_scalar)
    integer(c_int), parameter :: SX_POT_ENG = 0, SX_KIN_ENG = 1, SX_KIN_ROT = 2, SX_TEMP = 3, &
        SX_PRESSURE = 4, SX_DENSITY = 5, SX_VIRIAL = 6, SX_VIRIAL_BONDED = 7, SX_TARGET_TEMP = 8, SX_DOF = 9

    type, bind(C), public :: sx_step_spec
        integer(c_int) :: integrator, nforces
        integer(c_int) :: kinds(4)
        real(c_double) :: params(8, 4)
        integer(c_int) :: bonded
        real(c_double) :: dt, damping
        integer(c_int) :: first_step
    end type sx_step_spec

    interface
        integer(c_int) function sx_create(device, out) bind(C, name="sx_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: out
        end function
        integer(c_int) function sx_destroy(c) bind(C, name="sx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
        end function
        type(c_ptr) function sx_last_error(c) bind(C, name="sx_last_error")
            import :: c_ptr
            type(c_ptr), value :: c
        end function
        integer(c_int) function sx_set_num_atoms(c, num_atoms, rigid_atoms) bind(C, name="sx_set_num_atoms")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
            integer(c_int), value :: num_atoms, rigid_atoms
        end function
        integer(c_int) function sx_set_box(c, b, periodic) bind(C, name="sx_set_box")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), intent(in) :: b(3)
            integer(c_int), value :: periodic
        end function
        integer(c_int) function sx_upload(c, field, host, n) bind(C, name="sx_upload")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: field, n
            real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function sx_download(c, field, host, n) bind(C, name="sx_download")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: field, n
            real(c_double), intent(out) :: host(*)
        end function
        integer(c_int) function sx_set_scalar(c, id, v) bind(C, name="sx_set_scalar")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: id
            real(c_double), value :: v
        end function
        integer(c_int) function sx_get_scalar(c, id, v) bind(C, name="sx_get_scalar")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: id
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function sx_set_constraints(c, nc_, bbi1, bbj1, blen) bind(C, name="sx_set_constraints")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: nc_
            integer(c_int), intent(in) :: bbi1(*), bbj1(*)
            real(c_double), intent(in) :: blen(*)
        end function
        integer(c_int) function sx_initialize_forces(c) bind(C, name="sx_initialize_forces")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
        end function
        integer(c_int) function sx_calculate_thermo(c) bind(C, name="sx_calculate_thermo")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
        end function
        integer(c_int) function sx_nlist_init(c, skin, cutoff) bind(C, name="sx_nlist_init")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: skin, cutoff
        end function
        integer(c_int) function sx_nlist_update(c) bind(C, name="sx_nlist_update")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
        end function
        integer(c_int) function sx_forces(c, kind, params, nparams) bind(C, name="sx_forces")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: kind, nparams
            real(c_double), intent(in) :: params(*)
        end function
        integer(c_int) function sx_vv_initial_step(c, dt) bind(C, name="sx_vv_initial_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt
        end function
        integer(c_int) function sx_vv_final_step(c, dt) bind(C, name="sx_vv_final_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt
        end function
        integer(c_int) function sx_qq_initial_step(c, dt) bind(C, name="sx_qq_initial_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt
        end function
        integer(c_int) function sx_qq_final_step(c, dt) bind(C, name="sx_qq_final_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt
        end function
        integer(c_int) function sx_nht_initial_step(c, dt, coupling) bind(C, name="sx_nht_initial_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt
            type(c_ptr), value :: coupling
        end function
        integer(c_int) function sx_nht_final_step(c, dt, coupling) bind(C, name="sx_nht_final_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt
            type(c_ptr), value :: coupling
        end function
        integer(c_int) function sx_lgv_initial_step(c, dt, step, damping) bind(C, name="sx_lgv_initial_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt, damping
            integer(c_int), value :: step
        end function
        integer(c_int) function sx_lgv_final_step(c, dt, damping) bind(C, name="sx_lgv_final_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: dt, damping
        end function
        integer(c_int) function sx_run(c, spec, nsteps) bind(C, name="sx_run")
            import :: c_int, c_ptr, sx_step_spec
            type(c_ptr), value :: c
            type(sx_step_spec), intent(in) :: spec
            integer(c_int), value :: nsteps
        end function
    end interface

    public :: gpu_init, gpu_finalize, gpu_push_system, gpu_pull_system, gpu_init_nlist, gpu_update_nlist, &
        gpu_initialize_forces, gpu_calculate_forces, gpu_calculate_thermo, gpu_vv_initial_step, &
        gpu_vv_final_step, gpu_qq_initial_step, gpu_qq_final_step, gpu_nht_initial_step, &
        gpu_nht_final_step, gpu_lgv_initial_step, gpu_lgv_final_step, gpu_run

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: k
        if (rc == 0) return
        call c_f_pointer(sx_last_error(ctx), msg, [512])
        write(*, '(A, A, I0, A)', advance='no') what, ' failed (', rc, '): '
        do k = 1, 512
            if (msg(k) == c_null_char) exit
            write(*, '(A)', advance='no') msg(k)
        enddo
        write(*, *)
        error stop 1
    end subroutine check

    subroutine gpu_init(device)
        integer, intent(in) :: device
        call check(sx_create(int(device, c_int), ctx), 'sx_create')
    end subroutine gpu_init

    subroutine gpu_finalize()
        call check(sx_destroy(ctx), 'sx_destroy')
        ctx = c_null_ptr
    end subroutine gpu_finalize

    ! module system -> device (after set_num_atoms / initialize_system / the
    ! user's own set-up); the field order is enum sx_field
    subroutine gpu_push_system()
        integer(c_int) :: np
        integer(c_int), allocatable :: bi(:), bj(:)
        np = int(N, c_int)
        call check(sx_set_num_atoms(ctx, np, 1_c_int), 'sx_set_num_atoms')
        call check(sx_set_box(ctx, box, merge(1_c_int, 0_c_int, periodic)), 'sx_set_box')
        call put(0, RX); call put(1, RY); call put(2, RZ)
        call put(3, VX); call put(4, VY); call put(5, VZ)
        call put(6, FX); call put(7, FY); call put(8, FZ)
        call put(9, QW); call put(10, QX); call put(11, QY); call put(12, QZ)
        call put(13, LX); call put(14, LY); call put(15, LZ)
        call put(16, TX); call put(17, TY); call put(18, TZ)
        call put(19, mass); call put(20, inv_mass); call put(21, charge)
        call put(22, sig_x); call put(23, sig_y); call put(24, sig_z)
        call put(25, eps_x); call put(26, eps_y); call put(27, eps_z)
        call put(28, Ixx); call put(29, Iyy); call put(30, Izz)
        call put(31, Ixx_inv); call put(32, Iyy_inv); call put(33, Izz_inv)
        call check(sx_set_scalar(ctx, SX_POT_ENG, pot_eng), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_KIN_ENG, kin_eng), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_KIN_ROT, kin_rot), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_DENSITY, density), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_TARGET_TEMP, target_temp), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_DOF, real(dof, c_double)), 'sx_set_scalar')
        if (NC > 0) then
            bi = int(BBI(1:NC), c_int)
            bj = int(BBJ(1:NC), c_int)
            call check(sx_set_constraints(ctx, int(NC, c_int), bi, bj, BB_len), 'sx_set_constraints')
        endif
    contains
        subroutine put(field, a)
            integer, intent(in) :: field
            real(real64), intent(in) :: a(:)
            call check(sx_upload(ctx, int(field, c_int), a, np), 'sx_upload')
        end subroutine put
    end subroutine gpu_push_system

    ! device -> module system (for reporters and checkpoints)
    subroutine gpu_pull_system()
        integer(c_int) :: np
        np = int(N, c_int)
        call get(0, RX); call get(1, RY); call get(2, RZ)
        call get(3, VX); call get(4, VY); call get(5, VZ)
        call get(6, FX); call get(7, FY); call get(8, FZ)
        call get(9, QW); call get(10, QX); call get(11, QY); call get(12, QZ)
        call get(13, LX); call get(14, LY); call get(15, LZ)
        call get(16, TX); call get(17, TY); call get(18, TZ)
        call check(sx_get_scalar(ctx, SX_POT_ENG, pot_eng), 'sx_get_scalar')
        call check(sx_get_scalar(ctx, SX_KIN_ENG, kin_eng), 'sx_get_scalar')
        call check(sx_get_scalar(ctx, SX_KIN_ROT, kin_rot), 'sx_get_scalar')
        call check(sx_get_scalar(ctx, SX_TEMP, temp), 'sx_get_scalar')
        call check(sx_get_scalar(ctx, SX_PRESSURE, pressure), 'sx_get_scalar')
        call check(sx_get_scalar(ctx, SX_VIRIAL, virial), 'sx_get_scalar')
        call check(sx_get_scalar(ctx, SX_VIRIAL_BONDED, virial_bonded), 'sx_get_scalar')
    contains
        subroutine get(field, a)
            integer, intent(in) :: field
            real(real64), intent(inout) :: a(:)
            call check(sx_download(ctx, int(field, c_int), a, np), 'sx_download')
        end subroutine get
    end subroutine gpu_pull_system

    subroutine gpu_init_nlist(skin, cutoff)          ! init_nlist, system.f90:895
        real(real64), intent(in) :: skin, cutoff
        call check(sx_nlist_init(ctx, skin, cutoff), 'sx_nlist_init')
    end subroutine gpu_init_nlist

    subroutine gpu_update_nlist()                    ! if (needs_rebuild()) call build_nlist
        call check(sx_nlist_update(ctx), 'sx_nlist_update')
    end subroutine gpu_update_nlist

    subroutine gpu_initialize_forces()               ! initialize_forces, system.f90:423
        call check(sx_initialize_forces(ctx), 'sx_initialize_forces')
    end subroutine gpu_initialize_forces

    ! the README pair loop over the neighbour list for one `*_calculate_forces`;
    ! params as that routine's dummy arguments (see include/sucrose.h)
    subroutine gpu_calculate_forces(kind, params)
        integer, intent(in) :: kind
        real(real64), intent(in) :: params(:)
        call check(sx_forces(ctx, int(kind, c_int), params, int(size(params), c_int)), 'sx_forces')
    end subroutine gpu_calculate_forces

    subroutine gpu_calculate_thermo()                ! calculate_thermo, system.f90:725
        call check(sx_calculate_thermo(ctx), 'sx_calculate_thermo')
    end subroutine gpu_calculate_thermo

    subroutine gpu_vv_initial_step(dt)
        real(real64), intent(in) :: dt
        call check(sx_vv_initial_step(ctx, dt), 'sx_vv_initial_step')
    end subroutine gpu_vv_initial_step

    subroutine gpu_vv_final_step(dt)
        real(real64), intent(in) :: dt
        call check(sx_vv_final_step(ctx, dt), 'sx_vv_final_step')
    end subroutine gpu_vv_final_step

    subroutine gpu_qq_initial_step(dt)
        real(real64), intent(in) :: dt
        call check(sx_qq_initial_step(ctx, dt), 'sx_qq_initial_step')
    end subroutine gpu_qq_initial_step

    subroutine gpu_qq_final_step(dt)
        real(real64), intent(in) :: dt
        call check(sx_qq_final_step(ctx, dt), 'sx_qq_final_step')
    end subroutine gpu_qq_final_step

    subroutine gpu_nht_initial_step(dt, coupling)
        real(real64), intent(in) :: dt
        real(real64), intent(in), optional, target :: coupling
        type(c_ptr) :: cp
        cp = c_null_ptr
        if (present(coupling)) cp = c_loc(coupling)
        call check(sx_nht_initial_step(ctx, dt, cp), 'sx_nht_initial_step')
    end subroutine gpu_nht_initial_step

    subroutine gpu_nht_final_step(dt, coupling)
        real(real64), intent(in) :: dt
        real(real64), intent(in), optional, target :: coupling
        type(c_ptr) :: cp
        cp = c_null_ptr
        if (present(coupling)) cp = c_loc(coupling)
        call check(sx_nht_final_step(ctx, dt, cp), 'sx_nht_final_step')
    end subroutine gpu_nht_final_step

    subroutine gpu_lgv_initial_step(dt, step, damping)
        real(real64), intent(in) :: dt, damping
        integer, intent(in) :: step
        call check(sx_lgv_initial_step(ctx, dt, int(step, c_int), damping), 'sx_lgv_initial_step')
    end subroutine gpu_lgv_initial_step

    subroutine gpu_lgv_final_step(dt, damping)
        real(real64), intent(in) :: dt, damping
        call check(sx_lgv_final_step(ctx, dt, damping), 'sx_lgv_final_step')
    end subroutine gpu_lgv_final_step

    ! nsteps of the whole loop (initial step, list update, forces, final step,
    ! thermo) replayed as one captured HIP graph per step
    subroutine gpu_run(spec, nsteps)
        type(sx_step_spec), intent(in) :: spec
        integer, intent(in) :: nsteps
        call check(sx_run(ctx, spec, int(nsteps, c_int)), 'sx_run')
    end subroutine gpu_run

end module sucrose_gpu
