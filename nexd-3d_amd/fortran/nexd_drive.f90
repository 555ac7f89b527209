!--------------------------------------------------------------------------
! nexd_drive: a Fortran caller of libnexd_gpu.so through nexd_gpu_iface.
!
! Reads the operands of timeloop3D's time loop from a stream file (written
! by nexd_amd.fdump), hands them to the library exactly as a Fortran solver
! holding them would (column-major arrays, 1-based indices), runs all steps
! and writes the nine seismogram arrays (nrec,nt) to a stream file.
!
! usage: nexd_drive <operands file> <seismogram file>
!--------------------------------------------------------------------------
program nexd_drive
    use iso_c_binding
    use nexd_gpu_iface
    implicit none
    integer(c_int) :: hdr(12)
    integer :: nelem, np, npf, nt, nsrc, nrec, nglob, u, rc
    real(c_double) :: dt
    type(nexd_config) :: cfg
    type(c_ptr) :: ctx
    real(c_double), allocatable :: dr(:,:), ds(:,:), dtm(:,:), lift(:,:), mass(:,:)
    integer(c_int), allocatable :: fmask(:,:), ibool(:,:), neighbor(:,:), face(:,:), ibt(:,:,:), ibn(:,:,:)
    real(c_double), allocatable :: metric(:,:), jac(:), fscale(:,:), nn(:,:,:), st(:,:,:), tt(:,:,:)
    real(c_double), allocatable :: vpu(:), vsu(:), rho(:), lambdau(:), muu(:), mu(:), lambda(:)
    real(c_double), allocatable :: wl(:), anel(:), prof(:,:), srcarr(:,:,:,:), srcarrm(:,:,:,:), recint(:,:)
    integer(c_int), allocatable :: pmlcheck(:), srcelem(:), srctype(:), recelem(:)
    real(c_double), allocatable :: seis(:,:,:), energy(:)
    character(len=512) :: fin, fout
    integer :: k

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) hdr
    nelem = hdr(1); np = hdr(2); npf = hdr(3); nt = hdr(4); nsrc = hdr(5); nrec = hdr(6)
    cfg%nelem = nelem; cfg%order = hdr(7); cfg%nsrc = nsrc; cfg%nrec = nrec; cfg%nt = nt
    cfg%timeint = hdr(8); cfg%attenuation = hdr(9); cfg%pml = hdr(10); cfg%compat = NEXD_COMPAT_ALL
    cfg%track_displacement = 0; cfg%energy = 0; cfg%device = 0
    read(u) dt
    cfg%dt = dt
    nglob = nelem * np
    allocate(dr(np,np), ds(np,np), dtm(np,np), lift(np,4*npf), mass(np,np), fmask(npf,4))
    read(u) dr, ds, dtm, lift, mass, fmask
    allocate(ibool(np,nelem), metric(nglob,9), jac(nglob), fscale(4*npf,nelem))
    allocate(neighbor(4,nelem), face(4,nelem), ibt(npf,4,nelem), ibn(npf,4,nelem))
    allocate(nn(3,4,nelem), st(3,4,nelem), tt(3,4,nelem))
    read(u) ibool, metric, jac, fscale, neighbor, face, ibt, ibn, nn, st, tt
    allocate(vpu(nelem), vsu(nelem), rho(nelem), lambdau(nelem), muu(nelem), mu(nelem), lambda(nelem))
    read(u) vpu, vsu, rho, lambdau, muu, mu, lambda
    allocate(wl(3), anel(9*nelem), pmlcheck(nelem), prof(nglob,9))
    read(u) wl, anel, pmlcheck, prof
    allocate(srcelem(max(nsrc,1)), srctype(max(nsrc,1)), srcarr(np,3,max(nsrc,1),nt), srcarrm(np,6,max(nsrc,1),nt))
    read(u) srcelem(1:nsrc), srctype(1:nsrc), srcarr(:,:,1:nsrc,:), srcarrm(:,:,1:nsrc,:)
    allocate(recelem(max(nrec,1)), recint(np,max(nrec,1)))
    read(u) recelem(1:nrec), recint(:,1:nrec)
    close(u)

    rc = nexd_create(cfg, ctx)
    call check(rc, 'nexd_create', c_null_ptr)
    call check(nexd_upload_operators(ctx, dr, ds, dtm, lift, mass, fmask), 'upload_operators', ctx)
    call check(nexd_upload_mesh(ctx, ibool, metric(:,1), metric(:,2), metric(:,3), metric(:,4), metric(:,5), &
               metric(:,6), metric(:,7), metric(:,8), metric(:,9), jac, fscale, neighbor, face, ibt, ibn, &
               nn, st, tt), 'upload_mesh', ctx)
    call check(nexd_upload_material(ctx, vpu, vsu, rho, lambdau, muu, mu, lambda), 'upload_material', ctx)
    if (cfg%attenuation /= 0) call check(nexd_upload_anelastic(ctx, wl, anel), 'upload_anelastic', ctx)
    if (cfg%pml /= 0) call check(nexd_upload_pml(ctx, pmlcheck, prof(:,1), prof(:,2), prof(:,3), prof(:,4), &
               prof(:,5), prof(:,6), prof(:,7), prof(:,8), prof(:,9)), 'upload_pml', ctx)
    if (nsrc > 0) call check(nexd_upload_sources(ctx, srcelem, srctype, srcarr, srcarrm), 'upload_sources', ctx)
    if (nrec > 0) call check(nexd_upload_receivers(ctx, recelem, recint), 'upload_receivers', ctx)
    call check(nexd_reset(ctx), 'reset', ctx)
    allocate(energy(nt), seis(max(nrec,1),nt,9))
    call check(nexd_run(ctx, 1_c_int, int(nt, c_int), energy), 'run', ctx)
    call check(nexd_get_seismograms(ctx, seis(:,:,1), seis(:,:,2), seis(:,:,3), seis(:,:,4), seis(:,:,5), &
               seis(:,:,6), seis(:,:,7), seis(:,:,8), seis(:,:,9)), 'get_seismograms', ctx)
    call nexd_destroy(ctx)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    do k = 1, 9
        write(u) seis(1:nrec,:,k)
    end do
    close(u)
    write(*,*) 'nexd_drive: ', nt, ' steps done'

contains

    subroutine check(code, what, c)
        integer(c_int), intent(in) :: code
        character(len=*), intent(in) :: what
        type(c_ptr), intent(in) :: c
        if (code /= NEXD_OK) then
            write(*,*) 'nexd_drive: ', what, ' failed with ', code, ': ', trim(nexd_last_error(c))
            stop 1
        end if
    end subroutine check

end program nexd_drive
