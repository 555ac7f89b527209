!--------------------------------------------------------------------------
! The patch a maintainer applies to the reference's timeloop3D
! (src/timeloopMod.f90) so that libnexd_gpu.so runs its time loop.  Three
! fragments, each introduced by a marker line:
!   module        one line after the module's own use statements
!                 (after "use allocateMod", src/timeloopMod.f90:37)
!   declarations  before the first executable statement of timeloop3D
!                 (src/timeloopMod.f90:135)
!   loop          in place of the loop "do it=1,par%nt ... end do !timeloop"
!                 (src/timeloopMod.f90:357-589)
! The set-up before the loop and the seismogram and energy writers after it
! stay as they are.  oracle/splice_timeloop.py applies the fragments to a
! scratch copy of the reference (oracle/build_ref.sh <order> f64 gpu).
!
! What the loop fragment keeps from the reference's loop:
!   - the energy check after every step (src/timeloopMod.f90:517-541): the
!     library returns all_energy for every step; steps up to timecrit run
!     as one call, later ones one call per step while the check is armed;
!     when it fires, pmlcheck(:) = .false. and nexd_set_pml_active(0).  The
!     reference divides the window sums by locals avg_window1/avg_window2
!     that it never assigns (src/timeloopMod.f90:78,522,528); the fragment
!     divides by par%avg_window1/par%avg_window2;
!   - movie frames every par%frame steps (src/timeloopMod.f90:576-582): the
!     displacement comes back from nexd_get_state and the frame file is
!     written as the reference writes it;
!   - the run time in the reference's log line "simulation runs for".
!--------------------------------------------------------------------------
!>>> module
    use nexd_gpu_iface
!>>> declarations
        ! libnexd_gpu.so in place of the time loop
        type(nexd_config) :: gcfg
        type(c_ptr) :: gctx
        integer(c_int), allocatable :: gpml(:)
        real(c_double), allocatable :: genergy(:)
        real(c_double), allocatable, target :: gux(:), guy(:), guz(:)
        integer :: git, gend
!>>> loop
        gcfg%nelem = mesh%nelem
        gcfg%order = ORDER
        gcfg%nt = par%nt
        gcfg%timeint = par%timeint
        gcfg%nsrc = 0
        if (mesh%has_src) gcfg%nsrc = src%nsrc
        gcfg%nrec = 0
        if (mesh%has_rec) gcfg%nrec = rec%nrec
        gcfg%attenuation = merge(1, 0, par%attenuation)
        gcfg%pml = merge(1, 0, par%set_pml)
        gcfg%compat = NEXD_COMPAT_ALL
        gcfg%track_displacement = merge(1, 0, par%movie)
        gcfg%energy = gcfg%pml
        gcfg%device = 0
        gcfg%dt = deltat
        call nexd_check(c_null_ptr, nexd_create(gcfg, gctx), 'nexd_create')
        call nexd_check(gctx, nexd_upload_operators(gctx, mesh%Dr, mesh%Ds, mesh%Dt, mesh%lift, mesh%mass, &
                        mesh%fmask), 'nexd_upload_operators')
        call nexd_check(gctx, nexd_upload_mesh(gctx, mesh%ibool, mesh%rx, mesh%ry, mesh%rz, mesh%sx, mesh%sy, &
                        mesh%sz, mesh%tx, mesh%ty, mesh%tz, mesh%jacobian, mesh%fscale, mesh%neighbor, mesh%face, &
                        mesh%ibt, mesh%ibn, mesh%nnormal, mesh%stangential, mesh%ttangential), 'nexd_upload_mesh')
        call nexd_check(gctx, nexd_upload_material(gctx, mesh%vpu, mesh%vsu, mesh%rho, mesh%lambdau, mesh%muu, &
                        mesh%mu, mesh%lambda), 'nexd_upload_material')
        if (par%attenuation) call nexd_check(gctx, nexd_upload_anelastic(gctx, mesh%wl, anelasticvar), &
                                             'nexd_upload_anelastic')
        if (par%set_pml) then
            allocate(gpml(mesh%nelem))
            gpml = merge(1, 0, pmlcheck)
            call nexd_check(gctx, nexd_upload_pml(gctx, gpml, ddx, ddy, ddz, alphax, alphay, alphaz, kx, ky, kz), &
                            'nexd_upload_pml')
        end if
        if (mesh%has_src) call nexd_check(gctx, nexd_upload_sources(gctx, src%srcelem, src%srctype, srcArray, &
                                          srcArrayM), 'nexd_upload_sources')
        if (mesh%has_rec) call nexd_check(gctx, nexd_upload_receivers(gctx, rec%recelem, recInt), &
                                          'nexd_upload_receivers')
        call nexd_check(gctx, nexd_reset(gctx), 'nexd_reset')
        allocate(genergy(par%nt))
        genergy = 0.0
        if (par%movie) allocate(gux(size(ux)), guy(size(uy)), guz(size(uz)))
        git = 1
        do while (git <= par%nt)
            gend = par%nt
            if (par%set_pml .and. pmlcrit) gend = min(gend, max(git, timecrit))
            if (par%movie) gend = min(gend, ((git + par%frame - 1) / par%frame) * par%frame)
            call nexd_check(gctx, nexd_run(gctx, git, gend, genergy(git:gend)), 'nexd_run')
            if (par%set_pml) all_energy(git:gend) = genergy(git:gend)
            it = gend
            if (par%set_pml .and. pmlcrit .and. it > timecrit) then
                avg_energy1 = 0.0
                do i = 1, par%avg_window1
                    if (it - i > 0) avg_energy1 = avg_energy1 + all_energy(it - i) / par%avg_window1
                end do
                avg_energy2 = 0.0
                do i = 1, par%avg_window2
                    if (it - i > 0) avg_energy2 = avg_energy2 + all_energy(it - i) / par%avg_window2
                end do
                sta_lta = avg_energy2 / avg_energy1
                if (abs(sta_lta - 1) > par%sta_lta_trigger) then
                    write(*,*) ''
                    write(*,*) '!! Energy grows too much, pml switched off !!'
                    write(*,*) ''
                    pmlcrit = .false.
                    pmlcheck(:) = .false.
                    pmlswitch = it
                    call nexd_check(gctx, nexd_set_pml_active(gctx, 0_c_int), 'nexd_set_pml_active')
                end if
            end if
            if (par%movie .and. mod(it, par%frame) == 0) then
                call nexd_check(gctx, nexd_get_state(gctx, c_null_ptr, c_loc(gux), c_loc(guy), c_loc(guz)), &
                                'nexd_get_state')
                ux = gux
                uy = guy
                uz = guz
                uplot(:) = sqrt(ux(:)**2 + uy(:)**2 + uz(:)**2)
                write(filename, "('/moviedata',i6.6,'_it',i7.7,'.bin')") myrank + 1, it
                filename = trim(outpath)//trim(filename)
                open(unit=27, file=trim(filename), form="UNFORMATTED", status='unknown')
                write(27) uplot
                close(27)
            end if
            git = gend + 1
        end do
        if (mesh%has_rec) call nexd_check(gctx, nexd_get_seismograms(gctx, plotux, plotuy, plotuz, plotvx, &
                                          plotvy, plotvz, plotax, plotay, plotaz), 'nexd_get_seismograms')
        call nexd_destroy(gctx)
        if (par%log .and. myrank == 0) then
            call system_clock(t2, rate, count_max)
            dtges = real(t2 - t1, kind=8) / real(rate, kind=8)
            write(*,*) " simulation runs for ", dtges, " seconds"
            if (par%set_pml .and. pmlswitch > 0) write(*,*) 'PML was switched off at time step ', pmlswitch
        end if
!<<<
