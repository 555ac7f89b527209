!--------------------------------------------------------------------------
! nexd_gpu_iface: iso_c_binding interfaces of libnexd_gpu.so for a Fortran
! caller.  Arrays are passed as the reference holds them (column-major,
! 1-based index values, REAL(8) when CUSTOM_REAL = SIZE_DOUBLE); the library
! copies what it needs.  Each function returns 0 on success, a negative
! NEXD_E_* code otherwise (nexd_last_error gives the message).
!--------------------------------------------------------------------------
module nexd_gpu_iface
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: NEXD_OK = 0
    integer(c_int), parameter :: NEXD_COMPAT_RK3_FLOAT_WEIGHTS = 1
    integer(c_int), parameter :: NEXD_COMPAT_PML_LAST_ELEMENT = 2
    integer(c_int), parameter :: NEXD_COMPAT_ALL = 3

    type, bind(C) :: nexd_config
        integer(c_int) :: nelem, order, nsrc, nrec, nt, timeint
        integer(c_int) :: attenuation, pml, compat, track_displacement, energy, device
        real(c_double) :: dt
    end type nexd_config

    interface
        integer(c_int) function nexd_create(cfg, ctx) bind(C, name="nexd_create")
            import :: c_int, c_ptr, nexd_config
            type(nexd_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function nexd_upload_operators(ctx, dr, ds, dt, lift, mass, fmask) &
                bind(C, name="nexd_upload_operators")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: dr(*), ds(*), dt(*), lift(*), mass(*)
            integer(c_int), intent(in) :: fmask(*)
        end function

        integer(c_int) function nexd_upload_mesh(ctx, ibool, rx, ry, rz, sx, sy, sz, tx, ty, tz, &
                jacobian, fscale, neighbor, face, ibt, ibn, nnormal, stangential, ttangential) &
                bind(C, name="nexd_upload_mesh")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: ibool(*), neighbor(*), face(*), ibt(*), ibn(*)
            real(c_double), intent(in) :: rx(*), ry(*), rz(*), sx(*), sy(*), sz(*), tx(*), ty(*), tz(*)
            real(c_double), intent(in) :: jacobian(*), fscale(*), nnormal(*), stangential(*), ttangential(*)
        end function

        integer(c_int) function nexd_upload_material(ctx, vpu, vsu, rho, lambdau, muu, mu, lambda) &
                bind(C, name="nexd_upload_material")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: vpu(*), vsu(*), rho(*), lambdau(*), muu(*), mu(*), lambda(*)
        end function

        integer(c_int) function nexd_upload_anelastic(ctx, wl, anelasticvar) bind(C, name="nexd_upload_anelastic")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: wl(*), anelasticvar(*)
        end function

        integer(c_int) function nexd_upload_pml(ctx, pmlcheck, ddx, ddy, ddz, alphax, alphay, alphaz, &
                kx, ky, kz) bind(C, name="nexd_upload_pml")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: pmlcheck(*)
            real(c_double), intent(in) :: ddx(*), ddy(*), ddz(*), alphax(*), alphay(*), alphaz(*)
            real(c_double), intent(in) :: kx(*), ky(*), kz(*)
        end function

        integer(c_int) function nexd_upload_sources(ctx, srcelem, srctype, src_array, src_array_m) &
                bind(C, name="nexd_upload_sources")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: srcelem(*), srctype(*)
            real(c_double), intent(in) :: src_array(*), src_array_m(*)
        end function

        integer(c_int) function nexd_upload_receivers(ctx, recelem, rec_int) bind(C, name="nexd_upload_receivers")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: recelem(*)
            real(c_double), intent(in) :: rec_int(*)
        end function

        integer(c_int) function nexd_reset(ctx) bind(C, name="nexd_reset")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function nexd_run(ctx, it_begin, it_end, energy) bind(C, name="nexd_run")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: it_begin, it_end
            real(c_double), intent(out) :: energy(*)
        end function

        integer(c_int) function nexd_set_pml_active(ctx, active) bind(C, name="nexd_set_pml_active")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: active
        end function

        integer(c_int) function nexd_get_seismograms(ctx, ux, uy, uz, vx, vy, vz, ax, ay, az) &
                bind(C, name="nexd_get_seismograms")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: ux(*), uy(*), uz(*), vx(*), vy(*), vz(*), ax(*), ay(*), az(*)
        end function

        ! arrays passed as c_loc(x) or c_null_ptr (NULL: not wanted / keep)
        integer(c_int) function nexd_get_state(ctx, q, ux, uy, uz) bind(C, name="nexd_get_state")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, q, ux, uy, uz
        end function

        integer(c_int) function nexd_set_state(ctx, it, q, theta, fprime, gprime, hprime, ux, uy, uz) &
                bind(C, name="nexd_set_state")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
            type(c_ptr), value :: q, theta, fprime, gprime, hprime, ux, uy, uz
        end function

        subroutine nexd_destroy(ctx) bind(C, name="nexd_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine

        type(c_ptr) function nexd_last_error_c(ctx) bind(C, name="nexd_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
    end interface

contains

    ! message of the last failure as a Fortran string
    function nexd_last_error(ctx) result(msg)
        type(c_ptr), value :: ctx
        character(len=256) :: msg
        character(kind=c_char), pointer :: cs(:)
        type(c_ptr) :: p
        integer :: i
        msg = ' '
        p = nexd_last_error_c(ctx)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, cs, [256])
        do i = 1, 256
            if (cs(i) == c_null_char) exit
            msg(i:i) = cs(i)
        end do
    end function nexd_last_error

    ! stops the program with the library's message when rc is an error
    ! (the reference stops on errors too, src/mpiMod.f90:287-307)
    subroutine nexd_check(ctx, rc, what)
        type(c_ptr), value :: ctx
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= NEXD_OK) then
            write(*,*) 'libnexd_gpu: ', what, ' failed (', rc, '): ', trim(nexd_last_error(ctx))
            error stop 1
        end if
    end subroutine nexd_check

end module nexd_gpu_iface
