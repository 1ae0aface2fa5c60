! droplet_amd.f90 -- iso_c_binding interface to libdroplet_amd.so (include/droplet.h).
!
! Gives a Fortran driver written like the reference's `program main`
! (program/lbm_drop_xy.f90:756-868) the same call surface:
!   par + allocation      -> drp_create
!   ini + ini_op          -> drp_init_droplets / drp_init_fields
!   loop body             -> drp_step
!   output / outputfg     -> drp_get_macro / drp_get_populations
!   step_input injection  -> drp_inject_momentum
!   break_up3, diagnostics-> drp_deformation / drp_reduce
! Fields are returned in the reference's x-fastest order, i.e. a Fortran array
! phi(0:nx-1, 0:ny-1, 0:nz-1) receives them directly.
module droplet_amd
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: DRP_OK = 0
    integer(c_int), parameter :: DRP_FP64 = 0, DRP_FP32_SHIFTED = 1
    integer(c_int), parameter :: DRP_WALL_NONE = 0, DRP_WALL_Y = 2, DRP_WALL_Z = 3
    integer(c_int), parameter :: DRP_FORCE_NONE = 0, DRP_FORCE_TAYLOR_GREEN = 1, DRP_FORCE_FIELD = 2, &
                                 DRP_FORCE_SPECTRAL = 3
    integer(c_int), parameter :: DRP_FILE_FIELDS = 0, DRP_FILE_FG = 1, DRP_FILE_G = 2, DRP_FILE_FG_G = 3, &
                                 DRP_FILE_LAYOUT_XZ = 256
    integer(c_int), parameter :: DRP_GRAD_MOBILITY = 0, DRP_GRAD_KAPPA_G = 1, DRP_GRAD_WALL_U = 2
    integer(c_int), parameter :: DRP_FIELD_PHI = 0, DRP_FIELD_P = 1, DRP_FIELD_U1 = 2, DRP_FIELD_U2 = 3, &
                                 DRP_FIELD_U3 = 4
    integer(c_int), parameter :: DRP_POP_PHI = 0, DRP_POP_MOM = 1
    integer(c_int), parameter :: DRP_DERIVED_CHEMICAL_POTENTIAL = 0, DRP_DERIVED_FREE_ENERGY = 1, &
                                 DRP_DERIVED_ADVECTION = 2

    type, bind(C) :: drp_config
        integer(c_int32_t) :: nx, ny, nz, precision
        real(c_double) :: tau_f, nu1, nu2, kappa_f, kappa_g, phi1, phi2, a, b, T, anu
        integer(c_int32_t) :: wall_axis, wall_start
        real(c_double) :: wall_u
        integer(c_int32_t) :: force_kind, force_start
        real(c_double) :: force_amp, force_k, les_cs, les_delta
        integer(c_int32_t) :: les_ranks_x, les_ranks_y, les_ranks_z
        integer(c_int32_t) :: model, les_model
        integer(c_int32_t) :: advect_k_high, advect_k_low, advect_start
        real(c_double) :: ac_a, ac_width
    end type drp_config

    type, bind(C) :: drp_drop
        real(c_double) :: xc, yc, zc, diameter
    end type drp_drop

    type, bind(C) :: drp_stats
        real(c_double) :: phi_mass, phi_min, phi_max, kinetic, dissipation, interface_area, free_energy, p_mean
        real(c_double) :: momentum(3)
        real(c_double) :: interaction
    end type drp_stats

    interface
        integer(c_int) function drp_abi_sizes(config, drop, stats) bind(C, name="drp_abi_sizes")
            import :: c_int, c_size_t
            integer(c_size_t), intent(out) :: config, drop, stats
        end function
        integer(c_int) function drp_config_default(cfg) bind(C, name="drp_config_default")
            import :: c_int, drp_config
            type(drp_config), intent(inout) :: cfg
        end function
        integer(c_int) function drp_create(cfg, lat) bind(C, name="drp_create")
            import :: c_int, c_ptr, drp_config
            type(drp_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: lat
        end function
        integer(c_int) function drp_init_droplets(lat, drops, ndrops, p_init) bind(C, name="drp_init_droplets")
            import :: c_int, c_ptr, c_double, drp_drop
            type(c_ptr), value :: lat
            type(drp_drop), intent(in) :: drops(*)
            integer(c_int), value :: ndrops
            real(c_double), value :: p_init
        end function
        integer(c_int) function drp_init_fields(lat, phi, p, u1, u2, u3) bind(C, name="drp_init_fields")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            real(c_double), intent(in) :: phi(*), p(*), u1(*), u2(*), u3(*)
        end function
        integer(c_int) function drp_step(lat, nsteps) bind(C, name="drp_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: lat
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function drp_sync(lat) bind(C, name="drp_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: lat
        end function
        integer(c_int64_t) function drp_step_count(lat) bind(C, name="drp_step_count")
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: lat
        end function
        integer(c_int) function drp_get_macro(lat, field, host, n) bind(C, name="drp_get_macro")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            integer(c_int), value :: field
            real(c_double), intent(out) :: host(*)
            integer(c_size_t), value :: n
        end function
        integer(c_int) function drp_get_populations(lat, set, host, n) bind(C, name="drp_get_populations")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            integer(c_int), value :: set
            real(c_double), intent(out) :: host(*)
            integer(c_size_t), value :: n
        end function
        integer(c_int) function drp_set_populations(lat, set, host, n) bind(C, name="drp_set_populations")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            integer(c_int), value :: set
            real(c_double), intent(in) :: host(*)
            integer(c_size_t), value :: n
        end function
        integer(c_int) function drp_inject_momentum(lat, g, n) bind(C, name="drp_inject_momentum")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            real(c_double), intent(in) :: g(*)
            integer(c_size_t), value :: n
        end function
        integer(c_int) function drp_reduce(lat, stats) bind(C, name="drp_reduce")
            import :: c_int, c_ptr, drp_stats
            type(c_ptr), value :: lat
            type(drp_stats), intent(out) :: stats
        end function
        integer(c_int) function drp_reduce_noalloc(lat, stats) bind(C, name="drp_reduce_noalloc")
            import :: c_int, c_ptr, drp_stats
            type(c_ptr), value :: lat
            type(drp_stats), intent(out) :: stats
        end function
        integer(c_int) function drp_init_droplets_tanh(lat, drops, ndrops, p_init, width) &
                bind(C, name="drp_init_droplets_tanh")
            import :: c_int, c_ptr, c_double, drp_drop
            type(c_ptr), value :: lat
            type(drp_drop), intent(in) :: drops(*)
            integer(c_int), value :: ndrops
            real(c_double), value :: p_init, width
        end function
        ! enesupe_cal (turbulence_hit_xy.f90) and contribution_each_scale (drop_analysis.f90)
        integer(c_int) function drp_energy_spectrum(lat, E, nshells) bind(C, name="drp_energy_spectrum")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            real(c_double), intent(out) :: E(*)
            integer(c_int), value :: nshells
        end function
        integer(c_int) function drp_scale_interaction(lat, C, nshells) bind(C, name="drp_scale_interaction")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            real(c_double), intent(out) :: C(*)
            integer(c_int), value :: nshells
        end function
        integer(c_int) function drp_get_derived(lat, kind, out, n) bind(C, name="drp_get_derived")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            integer(c_int), value :: kind
            real(c_double), intent(out) :: out(*)
            integer(c_size_t), value :: n
        end function
        integer(c_int) function drp_phi_contribution(lat, C, nshells, bands) bind(C, name="drp_phi_contribution")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            real(c_double), intent(out) :: C(*)
            integer(c_int), value :: nshells
            real(c_double), intent(out) :: bands(4)
        end function
        integer(c_int) function drp_step_energy_budget(lat, terms) bind(C, name="drp_step_energy_budget")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            real(c_double), intent(out) :: terms(7)
        end function
        integer(c_int) function drp_deformation(lat, mode, L, B, D) bind(C, name="drp_deformation")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            integer(c_int), value :: mode
            real(c_double), intent(out) :: L, B, D
        end function
        ! forcing arrays of fg_cal / externalforce
        integer(c_int) function drp_set_force(lat, fx, fy, fz, n) bind(C, name="drp_set_force")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            real(c_double), intent(in) :: fx(*), fy(*), fz(*)
            integer(c_size_t), value :: n
        end function
        integer(c_int) function drp_get_force(lat, fx, fy, fz, n) bind(C, name="drp_get_force")
            import :: c_int, c_ptr, c_double, c_size_t
            type(c_ptr), value :: lat
            real(c_double), intent(out) :: fx(*), fy(*), fz(*)
            integer(c_size_t), value :: n
        end function
        ! droplet census of the current state (the count of program/percolation.f90):
        ! labels = c_null_ptr or c_loc of an integer(c_int32_t) array of n nodes
        integer(c_int) function drp_droplet_census(lat, threshold, count, volumes, first, cap, labels, n) &
                bind(C, name="drp_droplet_census")
            import :: c_int, c_ptr, c_double, c_int64_t, c_size_t
            type(c_ptr), value :: lat
            real(c_double), value :: threshold
            integer(c_int64_t), intent(out) :: count
            integer(c_int64_t), intent(out) :: volumes(*), first(*)
            integer(c_int64_t), value :: cap
            type(c_ptr), value :: labels
            integer(c_size_t), value :: n
        end function
        ! output / outputfg / step_input files in the reference's per-rank layout
        integer(c_int) function drp_write_reference_files_c(lat, dir, n, ranks_x, ranks_y, kind) &
                bind(C, name="drp_write_reference_files")
            import :: c_int, c_ptr, c_char, c_int64_t
            type(c_ptr), value :: lat
            character(kind=c_char), intent(in) :: dir(*)
            integer(c_int64_t), value :: n
            integer(c_int), value :: ranks_x, ranks_y, kind
        end function
        integer(c_int) function drp_read_reference_files_c(lat, dir, n, ranks_x, ranks_y, kind) &
                bind(C, name="drp_read_reference_files")
            import :: c_int, c_ptr, c_char, c_int64_t
            type(c_ptr), value :: lat
            character(kind=c_char), intent(in) :: dir(*)
            integer(c_int64_t), value :: n
            integer(c_int), value :: ranks_x, ranks_y, kind
        end function
        ! timing and tile shape
        integer(c_int) function drp_time_steps(lat, nsteps, ms) bind(C, name="drp_time_steps")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: lat
            integer(c_int), value :: nsteps
            real(c_double), intent(out) :: ms
        end function
        integer(c_int) function drp_set_tile_shape(lat, shape) bind(C, name="drp_set_tile_shape")
            import :: c_int, c_ptr
            type(c_ptr), value :: lat
            integer(c_int), value :: shape
        end function
        ! ensembles: one launch per step over all members
        integer(c_int) function drp_batch_create(cfgs, count, batch) bind(C, name="drp_batch_create")
            import :: c_int, c_ptr, drp_config
            type(drp_config), intent(in) :: cfgs(*)
            integer(c_int), value :: count
            type(c_ptr), intent(out) :: batch
        end function
        type(c_ptr) function drp_batch_member(batch, i) bind(C, name="drp_batch_member")
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: i
        end function
        integer(c_int) function drp_batch_step(batch, nsteps) bind(C, name="drp_batch_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function drp_batch_sync(batch) bind(C, name="drp_batch_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function
        subroutine drp_batch_destroy(batch) bind(C, name="drp_batch_destroy")
            import :: c_ptr
            type(c_ptr), value :: batch
        end subroutine
        ! parameter sensitivities (checkpointed adjoint)
        integer(c_int) function drp_tape_create(lat, checkpoint_every, tape) bind(C, name="drp_tape_create")
            import :: c_int, c_ptr
            type(c_ptr), value :: lat
            integer(c_int), value :: checkpoint_every
            type(c_ptr), intent(out) :: tape
        end function
        integer(c_int) function drp_tape_step(tape, nsteps) bind(C, name="drp_tape_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: tape
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function drp_tape_backward(tape, w, loss, grad) bind(C, name="drp_tape_backward")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: tape
            real(c_double), intent(in) :: w(2)
            real(c_double), intent(out) :: loss, grad(3)
        end function
        subroutine drp_tape_destroy(tape) bind(C, name="drp_tape_destroy")
            import :: c_ptr
            type(c_ptr), value :: tape
        end subroutine
        type(c_ptr) function drp_last_error_c() bind(C, name="drp_last_error")
            import :: c_ptr
        end function
        subroutine drp_destroy(lat) bind(C, name="drp_destroy")
            import :: c_ptr
            type(c_ptr), value :: lat
        end subroutine
    end interface

contains

    ! Message of the last failure as a Fortran string.
    function drp_last_error() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: k
        call c_f_pointer(drp_last_error_c(), s, [512])
        k = 0
        do while (k < 512)
            if (s(k + 1) == c_null_char) exit
            k = k + 1
        end do
        allocate(character(len=k) :: msg)
        msg = transfer(s(1:k), msg)
    end function

    ! Stop with the library's message when rc is not DRP_OK.
    ! Fortran-string wrappers of the file entry points (NUL-terminated for C)
    integer(c_int) function drp_write_reference_files(lat, dir, n, ranks_x, ranks_y, kind)
        type(c_ptr), intent(in) :: lat
        character(len=*), intent(in) :: dir
        integer(c_int64_t), intent(in) :: n
        integer(c_int), intent(in) :: ranks_x, ranks_y, kind
        drp_write_reference_files = drp_write_reference_files_c(lat, trim(dir)//c_null_char, n, ranks_x, ranks_y, kind)
    end function

    integer(c_int) function drp_read_reference_files(lat, dir, n, ranks_x, ranks_y, kind)
        type(c_ptr), intent(in) :: lat
        character(len=*), intent(in) :: dir
        integer(c_int64_t), intent(in) :: n
        integer(c_int), intent(in) :: ranks_x, ranks_y, kind
        drp_read_reference_files = drp_read_reference_files_c(lat, trim(dir)//c_null_char, n, ranks_x, ranks_y, kind)
    end function

    subroutine drp_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: where
        if (rc /= DRP_OK) then
            write(*, '(a, a, i0, a, a)') where, ' failed (', rc, '): ', drp_last_error()
            error stop 1
        end if
    end subroutine

    ! Check that this module's struct layouts match the compiled library.
    logical function drp_layout_ok()
        integer(c_size_t) :: sc, sd, ss
        type(drp_config) :: c
        type(drp_drop) :: d
        type(drp_stats) :: s
        integer(c_int) :: v
        v = drp_abi_sizes(sc, sd, ss)
        drp_layout_ok = (v == 3) .and. (sc == c_sizeof(c)) .and. (sd == c_sizeof(d)) .and. (ss == c_sizeof(s))
    end function

end module droplet_amd
