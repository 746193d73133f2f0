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
        SX_GB = 6, SX_GBR = 7, SX_ECP = 8, SX_ECPR = 9, SX_DH = 10, SX_GBW = 11, SX_GBC = 12, SX_GBWC = 13, &
        SX_ECPW = 14, SX_ECPC = 15, SX_ECPWC = 16, SX_DIPOLE = 17, SX_COULOMB = 18
    ! reporters (sx_report_kind)
    integer(c_int), parameter, public :: SX_REPORT_STATE_DATA = 1, SX_REPORT_TRAJECTORY = 2, &
        SX_REPORT_CONFIG = 3, SX_REPORT_CHECKPOINT = 4
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
        integer(c_int) function sx_get_box(c, b) bind(C, name="sx_get_box")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), intent(out) :: b(3)
        end function
        integer(c_int) function sx_set_atom_types(c, rt) bind(C, name="sx_set_atom_types")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
            integer(c_int), intent(in) :: rt(*)
        end function
        integer(c_int) function sx_set_target_pressure(c, p) bind(C, name="sx_set_target_pressure")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            real(c_double), value :: p
        end function
        integer(c_int) function sx_mc_barostat(c, spec, dv, u, acc) bind(C, name="sx_mc_barostat")
            import :: c_int, c_ptr, c_double, sx_step_spec
            type(c_ptr), value :: c
            type(sx_step_spec), intent(in) :: spec
            real(c_double), value :: dv
            real(c_double), intent(in) :: u(2)
            integer(c_int), intent(out) :: acc
        end function
        integer(c_int) function sx_reporter_open(c, kind, fname, r) bind(C, name="sx_reporter_open")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: c
            integer(c_int), value :: kind
            character(kind=c_char), intent(in) :: fname(*)
            type(c_ptr), intent(out) :: r
        end function
        integer(c_int) function sx_report(r, step) bind(C, name="sx_report")
            import :: c_int, c_ptr
            type(c_ptr), value :: r
            integer(c_int), value :: step
        end function
        integer(c_int) function sx_reporter_close(r) bind(C, name="sx_reporter_close")
            import :: c_int, c_ptr
            type(c_ptr), value :: r
        end function
        integer(c_int) function sx_load_checkpoint(c, fname, cfg) bind(C, name="sx_load_checkpoint")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: c
            character(kind=c_char), intent(in) :: fname(*)
            integer(c_int), intent(out) :: cfg
        end function
        integer(c_int) function sx_copy_transform(c, cnt, f1, t1, off, si) bind(C, name="sx_copy_transform")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: cnt, si
            integer(c_int), intent(in) :: f1(*), t1(*)
            real(c_double), intent(in) :: off(*)
        end function
        integer(c_int) function sx_transfer_forces(c, cnt, f1, t1) bind(C, name="sx_transfer_forces")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
            integer(c_int), value :: cnt
            integer(c_int), intent(in) :: f1(*), t1(*)
        end function
        integer(c_int) function sx_init_reaction_field(c) bind(C, name="sx_init_reaction_field")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
        end function
        integer(c_int) function sx_eval_reaction_field(c) bind(C, name="sx_eval_reaction_field")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
        end function
        integer(c_int) function sx_set_bonded(c, kind, cnt, idx1, coeff, angle) bind(C, name="sx_set_bonded")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: c
            integer(c_int), value :: kind, cnt
            integer(c_int), intent(in) :: idx1(*)
            real(c_double), value :: coeff, angle
        end function
        integer(c_int) function sx_bonded_forces(c) bind(C, name="sx_bonded_forces")
            import :: c_int, c_ptr
            type(c_ptr), value :: c
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
        gpu_nht_final_step, gpu_lgv_initial_step, gpu_lgv_final_step, gpu_run, gpu_set_target_pressure, &
        gpu_mc_barostat, gpu_reporter_open, gpu_report, gpu_reporter_close, gpu_load_checkpoint, &
        gpu_copy_transform, gpu_transfer_forces, gpu_init_reaction_field, gpu_eval_reaction_field, &
        gpu_set_bonded, gpu_bonded_forces

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
        call check(sx_set_scalar(ctx, SX_TEMP, temp), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_PRESSURE, pressure), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_DENSITY, density), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_TARGET_TEMP, target_temp), 'sx_set_scalar')
        call check(sx_set_scalar(ctx, SX_DOF, real(dof, c_double)), 'sx_set_scalar')
        call check(sx_set_atom_types(ctx, int(rtype, c_int)), 'sx_set_atom_types')
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
        call check(sx_get_scalar(ctx, SX_DENSITY, density), 'sx_get_scalar')
        call check(sx_get_box(ctx, box), 'sx_get_box')
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

    subroutine gpu_set_target_pressure(pres)         ! set_target_pressure, system.f90:224
        real(real64), intent(in) :: pres
        call check(sx_set_target_pressure(ctx, pres), 'sx_set_target_pressure')
    end subroutine gpu_set_target_pressure

    ! mc_barostat (integrators.f90:864-945): the two uniform numbers come from
    ! random_number in the order the reference draws them (the volume move,
    ! then the acceptance test); compute_potential is the list rebuilt at the
    ! current box plus the pair loops (and bonded terms) of spec
    subroutine gpu_mc_barostat(spec, accept, volume_change)
        type(sx_step_spec), intent(in) :: spec
        logical, intent(out) :: accept
        real(real64), intent(in), optional :: volume_change
        real(c_double) :: u(2), dv
        integer(c_int) :: acc
        call random_number(u(1))
        call random_number(u(2))
        dv = -1.0_c_double
        if (present(volume_change)) dv = volume_change
        call check(sx_mc_barostat(ctx, spec, dv, u, acc), 'sx_mc_barostat')
        accept = acc /= 0
    end subroutine gpu_mc_barostat

    ! open_*_file / report_* / close_*_file (reporters.f90) on the device
    ! state: a handle per open file
    function gpu_reporter_open(kind, filename) result(r)
        integer, intent(in) :: kind
        character(*), intent(in) :: filename
        type(c_ptr) :: r
        call check(sx_reporter_open(ctx, int(kind, c_int), trim(filename) // c_null_char, r), 'sx_reporter_open')
    end function gpu_reporter_open

    subroutine gpu_report(r, step)
        type(c_ptr), intent(in) :: r
        integer, intent(in) :: step
        call check(sx_report(r, int(step, c_int)), 'sx_report')
    end subroutine gpu_report

    subroutine gpu_reporter_close(r)
        type(c_ptr), intent(inout) :: r
        call check(sx_reporter_close(r), 'sx_reporter_close')
        r = c_null_ptr
    end subroutine gpu_reporter_close

    subroutine gpu_load_checkpoint(filename, config_index)   ! load_checkpoint, reporters.f90:188
        character(*), intent(in) :: filename
        integer, intent(out), optional :: config_index
        integer(c_int) :: cfg
        call check(sx_load_checkpoint(ctx, trim(filename) // c_null_char, cfg), 'sx_load_checkpoint')
        if (present(config_index)) config_index = cfg
    end subroutine gpu_load_checkpoint

    subroutine gpu_copy_transform(from, to, xoffset, yoffset, zoffset, set_inertia)   ! system.f90:743
        integer, intent(in) :: from, to
        real(real64), intent(in) :: xoffset, yoffset, zoffset
        logical, intent(in), optional :: set_inertia
        integer(c_int) :: si
        si = -1
        if (present(set_inertia)) si = merge(1_c_int, 0_c_int, set_inertia)
        call check(sx_copy_transform(ctx, 1_c_int, [int(from, c_int)], [int(to, c_int)], &
                                     [xoffset, yoffset, zoffset], si), 'sx_copy_transform')
    end subroutine gpu_copy_transform

    subroutine gpu_transfer_forces(from, to)         ! transfer_forces, system.f90:844
        integer, intent(in) :: from, to
        call check(sx_transfer_forces(ctx, 1_c_int, [int(from, c_int)], [int(to, c_int)]), 'sx_transfer_forces')
    end subroutine gpu_transfer_forces

    subroutine gpu_init_reaction_field()             ! init_reaction_field, force.f90:2677
        call check(sx_init_reaction_field(ctx), 'sx_init_reaction_field')
    end subroutine gpu_init_reaction_field

    subroutine gpu_eval_reaction_field()             ! eval_reaction_field, force.f90:2753
        call check(sx_eval_reaction_field(ctx), 'sx_eval_reaction_field')
    end subroutine gpu_eval_reaction_field

    ! bonded terms registered once (kind as sx_set_bonded: 1 bend_, 2
    ! cosine_bend_, 3 dih_, 4 cosine_dih_, 5 angle_bend_, 6 angle_dih_);
    ! idx(:, t) holds term t's atoms
    subroutine gpu_set_bonded(kind, idx, coeff, angle)
        integer, intent(in) :: kind
        integer, intent(in) :: idx(:, :)
        real(real64), intent(in) :: coeff
        real(real64), intent(in), optional :: angle
        real(c_double) :: a
        a = 0.0_c_double
        if (present(angle)) a = angle
        call check(sx_set_bonded(ctx, int(kind, c_int), int(size(idx, 2), c_int), int(reshape(idx, [size(idx)]), c_int), &
                                 coeff, a), 'sx_set_bonded')
    end subroutine gpu_set_bonded

    subroutine gpu_bonded_forces()
        call check(sx_bonded_forces(ctx), 'sx_bonded_forces')
    end subroutine gpu_bonded_forces

end module sucrose_gpu
