! shear_driver.f90 -- a Fortran driver in the shape of the reference's shear-droplet
! program (prototype_program/two-phase/dropch.f90:1079-1912), running the step
! on the GPU through the droplet_amd iso_c_binding module.
!
!   shear_driver [N [nsteps [every]]]      (defaults 64 2000 10)
!
! Walls normal to y move -U/+U along x from step 1; one sharp-step droplet of
! radius N/5 at the centre.  Every `every` steps the Taylor deformation (break_up3
! and second-moment variants) is appended to deformation.d; at the end the fields
! are written as one Fortran unformatted record (phi, u1, u2, u3, p) per node in
! z, y, x loop order -- the record layout of the reference's `output`
! (program/lbm_drop_xy.f90:695-719) for a single rank.
program shear_driver
    use droplet_amd
    implicit none
    type(drp_config) :: cfg
    type(drp_drop) :: drop(1)
    type(drp_stats) :: st
    type(c_ptr) :: lat
    integer :: ng, nsteps, every, n, xi, yi, zi, nargs
    character(len=32) :: arg
    real(c_double) :: L1, B1, D1, L2, B2, D2, sigma
    real(c_double), allocatable :: phi(:,:,:), p(:,:,:), u1(:,:,:), u2(:,:,:), u3(:,:,:)
    integer(c_size_t) :: nn

    ng = 64; nsteps = 2000; every = 10
    nargs = command_argument_count()
    if (nargs >= 1) then; call get_command_argument(1, arg); read(arg, *) ng; end if
    if (nargs >= 2) then; call get_command_argument(2, arg); read(arg, *) nsteps; end if
    if (nargs >= 3) then; call get_command_argument(3, arg); read(arg, *) every; end if

    if (.not. drp_layout_ok()) then
        write(*, *) 'struct layout of the Fortran module does not match libdroplet_amd.so'
        error stop 2
    end if

    ! ---- par: reference defaults, then this case's parameters
    call drp_check(drp_config_default(cfg), 'drp_config_default')
    cfg%nx = ng; cfg%ny = ng; cfg%nz = ng
    cfg%tau_f = 0.7d0
    cfg%nu1 = 0.1d0; cfg%nu2 = 0.1d0
    cfg%kappa_f = 0.01d0
    sigma = 0.01d0
    cfg%kappa_g = (sigma / 1.7039d0)**(1.0d0 / 0.9991d0)
    cfg%wall_axis = DRP_WALL_Y; cfg%wall_start = 1; cfg%wall_u = 0.01d0
    call drp_check(drp_create(cfg, lat), 'drp_create')

    ! ---- ini + ini_op
    drop(1) = drp_drop(0.5d0 * dble(ng - 1), 0.5d0 * dble(ng - 1), 0.5d0 * dble(ng - 1), 0.4d0 * dble(ng))
    call drp_check(drp_init_droplets(lat, drop, 1, 1.0d0 / 3.0d0), 'drp_init_droplets')

    ! ---- time loop
    open(20, file='deformation.d', status='replace')
    do n = 1, nsteps
        call drp_check(drp_step(lat, 1), 'drp_step')
        if (mod(n, every) == 0) then
            call drp_check(drp_deformation(lat, 0, L1, B1, D1), 'drp_deformation')
            call drp_check(drp_deformation(lat, 2, L2, B2, D2), 'drp_deformation')
            write(20, '(i8, 6es24.16)') n, L1, B1, D1, L2, B2, D2
        end if
    end do
    close(20)

    call drp_check(drp_reduce(lat, st), 'drp_reduce')
    write(*, '(a, i0, a, es24.16, a, es24.16)') 'steps ', drp_step_count(lat), ' phi_mass ', st%phi_mass, &
        ' kinetic ', st%kinetic

    ! ---- output
    allocate(phi(0:ng-1, 0:ng-1, 0:ng-1), p(0:ng-1, 0:ng-1, 0:ng-1), u1(0:ng-1, 0:ng-1, 0:ng-1))
    allocate(u2(0:ng-1, 0:ng-1, 0:ng-1), u3(0:ng-1, 0:ng-1, 0:ng-1))
    nn = int(ng, c_size_t)**3
    call drp_check(drp_get_macro(lat, DRP_FIELD_PHI, phi, nn), 'drp_get_macro')
    call drp_check(drp_get_macro(lat, DRP_FIELD_P, p, nn), 'drp_get_macro')
    call drp_check(drp_get_macro(lat, DRP_FIELD_U1, u1, nn), 'drp_get_macro')
    call drp_check(drp_get_macro(lat, DRP_FIELD_U2, u2, nn), 'drp_get_macro')
    call drp_check(drp_get_macro(lat, DRP_FIELD_U3, u3, nn), 'drp_get_macro')
    open(21, file='fields.bin', form='unformatted', status='replace')
    do zi = 0, ng - 1
        do yi = 0, ng - 1
            do xi = 0, ng - 1
                write(21) phi(xi, yi, zi), u1(xi, yi, zi), u2(xi, yi, zi), u3(xi, yi, zi), p(xi, yi, zi)
            end do
        end do
    end do
    close(21)
    call drp_destroy(lat)
end program shear_driver
