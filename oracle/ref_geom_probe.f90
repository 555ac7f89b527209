! Test infrastructure only: never linked into the product library.
!
! Probe program linked against the reference's own Fortran modules (built by
! oracle/build_ref.sh from /root/reference/src).  For a tetrahedral mesh read
! from <infile> it calls the reference's geometry, location and attenuation
! routines and dumps what they return, so that the host-side precompute of
! this project can be compared with them:
!   - geometricFactors3d (src/geometricFactorsMod.f90:28-68) and normals3d
!     (src/normalsMod.f90:27-130) on selected elements, with the node
!     coordinates formed as the mesher forms them (src/meshMod.f90:812-827);
!   - tet_mesh_neighbor_tets (src/tetraNeighborMod.f90:563-763) on the whole
!     mesh, with the mesher's face reordering (src/meshMod.f90:253-258) and
!     its face-connection loop (src/meshMod.f90:519-532);
!   - findSource and findReceiver (src/sourceReceiverMod.f90:204-528);
!   - anelasticPreCalc and calcAnelasticCoefficients
!     (src/anelasticMod.f90:27-179) for a list of materials.
! The mesher computes a few quantities inline in createMesh, which needs
! METIS (absent) and cannot be called; the probe carries those loops over
! as they stand there and says so at each:
!   - fscale = sj / J at the face nodes (src/meshMod.f90:878-889);
!   - the face frames nnormal, stangential, ttangential
!     (src/meshMod.f90:843-876);
!   - ibt/ibn and the face-node permutation (src/meshMod.f90:939-987);
!   - dtfactor and minGLL_jac (src/meshMod.f90:1134-1157).
!
! Input (stream, little-endian): int32 nelem, ncoord, nsrc, nrec, nsel,
! nmat; real64 coord(3,ncoord); int32 elem(4,nelem) (1-based); real64
! srcxyz(3,nsrc), recxyz(3,nrec); int32 sel(nsel) (1-based); real64
! vpu(nelem); real64 f0_att, fr_att; real64 mat(5,nmat) = qp, qs, vp, vs, rho.
! Output (stream): see the write statements at the end; reals as real64.
!
! Usage: ref_geom_probe <infile> <outfile>
program ref_geom_probe
    use constantsMod
    use typeMod
    use nodesMod
    use tetTrafoMod
    use vandermondeMod
    use dmatricesMod
    use geometricFactorsMod
    use normalsMod
    use sourceReceiverMod
    use anelasticMod
    use tetraNeighborMod
    implicit none
    character(len=256) :: infile, outfile
    integer(4) :: nelem, ncoord, nsrc, nrec, nsel, nmat
    real(8), allocatable :: coord8(:,:), src8(:,:), rec8(:,:), vpu8(:), mat8(:,:)
    integer(4), allocatable :: elem(:,:), sel(:)
    real(8) :: f0_8, fr_8
    real(kind=CUSTOM_REAL), allocatable :: coord(:,:), vx(:), vy(:), vz(:), vpu(:)
    integer, allocatable :: ibool(:,:), neighbortemp(:,:), neighbor(:,:), face(:,:)
    integer, allocatable :: ibt(:,:,:), ibn(:,:,:)
    real(kind=CUSTOM_REAL), allocatable :: jac_all(:,:), sj_all(:,:)
    real(kind=CUSTOM_REAL), dimension(Np) :: x, y, z, r, s, t
    real(kind=CUSTOM_REAL), dimension(Np,Np) :: vdm, dr, ds, dt
    real(kind=CUSTOM_REAL), dimension(Np) :: rx, sx, tx, ry, sy, ty, rz, sz, tz, jac
    real(kind=CUSTOM_REAL), dimension(4*NpF) :: nx, ny, nz, sj, fscale
    real(kind=CUSTOM_REAL), dimension(3,4) :: nnormal, stangential, ttangential
    integer, dimension(NpF,4) :: fmask
    integer, dimension(NpF) :: perm
    integer, dimension(Np) :: iv, ivn
    real(kind=CUSTOM_REAL) :: rtemp1, rtemp2, rtemp3, rtemp4, min1, min2, minGLL, epsil_e
    real(kind=CUSTOM_REAL) :: f0_att, fr_att, lambda_u, mu_u, qp, qs, vp, vs, rho
    real(kind=CUSTOM_REAL), dimension(nMB) :: wl, ylambda, ymu
    real(kind=CUSTOM_REAL), dimension(2*nMB-1) :: wk
    type(parameterVar) :: par
    type(srcVar) :: src
    type(recVar) :: rec
    integer :: i, j, k, c, d, e, ie, is, in, l, m, nglob

    call get_command_argument(1, infile)
    call get_command_argument(2, outfile)
    open(unit=31, file=trim(infile), form='unformatted', access='stream', status='old')
    read(31) nelem, ncoord, nsrc, nrec, nsel, nmat
    allocate(coord8(3,ncoord), elem(4,nelem), src8(3,nsrc), rec8(3,nrec), sel(nsel), vpu8(nelem), mat8(5,nmat))
    read(31) coord8, elem, src8, rec8, sel, vpu8, f0_8, fr_8, mat8
    close(31)
    allocate(coord(3,ncoord), vpu(nelem))
    coord = real(coord8, CUSTOM_REAL)
    vpu = real(vpu8, CUSTOM_REAL)

    ! reference element and operators as the mesher sets them up
    ! (src/meshMod.f90:601-602: balpha(order); findSource and findReceiver
    ! build their own search nodes with balpha(NGLL))
    call nodes3D(balpha(order), x, y, z)
    call xyzToRst(x, y, z, r, s, t)
    call vdm3d(vdm, r, s, t)
    call dmatrices3d(dr, ds, dt, r, s, t, vdm)
    ! the four fmask loops of src/meshMod.f90:625-660
    fmask = 0
    d = 1
    do i = 1, Np
        if (abs(t(i)+1.0) < EPS) then
            fmask(d,1) = i; d = d+1
        end if
    end do
    d = 1
    do i = 1, Np
        if (abs(s(i)+1.0) < EPS) then
            fmask(d,2) = i; d = d+1
        end if
    end do
    d = 1
    do i = 1, Np
        if (abs(1.0+r(i)+s(i)+t(i)) < EPS) then
            fmask(d,3) = i; d = d+1
        end if
    end do
    d = 1
    do i = 1, Np
        if (abs(1.0+r(i)) < EPS) then
            fmask(d,4) = i; d = d+1
        end if
    end do
    minGLL = abs(r(1)-r(2))

    ! node coordinates and ibool (src/meshMod.f90:812-827)
    nglob = nelem*Np
    allocate(vx(nglob), vy(nglob), vz(nglob), ibool(Np,nelem))
    c = 0
    do i = 1, nelem
        do j = 1, Np
            c = c+1
            vx(c) = 0.5 * ( -(1.0+r(j)+s(j)+t(j)) * coord(1,elem(1,i)) + (1.0+r(j)) * coord(1,elem(2,i)) + &
                (1+s(j)) * coord(1,elem(3,i)) + (1+t(j)) * coord(1,elem(4,i)))
            vy(c) = 0.5 * ( -(1.0+r(j)+s(j)+t(j)) * coord(2,elem(1,i)) + (1.0+r(j)) * coord(2,elem(2,i)) + &
                (1+s(j)) * coord(2,elem(3,i)) + (1+t(j)) * coord(2,elem(4,i)))
            vz(c) = 0.5 * ( -(1.0+r(j)+s(j)+t(j)) * coord(3,elem(1,i)) + (1.0+r(j)) * coord(3,elem(2,i)) + &
                (1+s(j)) * coord(3,elem(3,i)) + (1+t(j)) * coord(3,elem(4,i)))
            ibool(j,i) = c
        end do
    end do

    ! neighbours: the reference routine, then the mesher's reordering and
    ! face connections
    allocate(neighbortemp(4,nelem), neighbor(4,nelem), face(4,nelem))
    call tet_mesh_neighbor_tets(4, nelem, elem, neighbortemp)
    neighbor(1,:) = neighbortemp(4,:)
    neighbor(2,:) = neighbortemp(3,:)
    neighbor(3,:) = neighbortemp(1,:)
    neighbor(4,:) = neighbortemp(2,:)
    where (neighbor <= 0) neighbor = 0
    face = 0
    do i = 1, nelem
        do j = 1, 4
            c = neighbor(j,i)
            if (c > 0) then
                do k = 1, 4
                    if (neighbor(k,c) == i) face(j,i) = k
                end do
            end if
        end do
    end do

    ! jacobian and sj of every element (needed by dtfactor)
    allocate(jac_all(Np,nelem), sj_all(4*NpF,nelem))
    do ie = 1, nelem
        iv = ibool(:,ie)
        call geometricFactors3d(rx, sx, tx, ry, sy, ty, rz, sz, tz, jac, vx(iv), vy(iv), vz(iv), dr, ds, dt)
        jac_all(:,ie) = jac
        call normals3d(nx, ny, nz, sj, dr, ds, dt, vx(iv), vy(iv), vz(iv), fmask)
        sj_all(:,ie) = sj
    end do

    ! ibt/ibn and the face-node permutation (src/meshMod.f90:939-987); the
    ! tolerance epsil is 1/180 of the element's summed edge lengths
    ! (src/meshMod.f90:366-378)
    allocate(ibt(NpF,4,nelem), ibn(NpF,4,nelem))
    ibt = 0
    ibn = 0
    do ie = 1, nelem
        iv = ibool(:,ie)
        do is = 1, 4
            in = neighbor(is,ie)
            ibt(:,is,ie) = iv(fmask(:,is))
            if (in > 0) then
                ivn = ibool(:,in)
                ibn(:,is,ie) = ivn(fmask(:,face(is,ie)))
            end if
        end do
    end do
    do ie = 1, nelem
        rtemp1 = 0.0
        do j = 1, 4
            do k = j+1, 4
                rtemp1 = rtemp1 + sqrt((coord(1,elem(j,ie))-coord(1,elem(k,ie)))**2 + &
                    (coord(2,elem(j,ie))-coord(2,elem(k,ie)))**2 + (coord(3,elem(j,ie))-coord(3,elem(k,ie)))**2)
            end do
        end do
        epsil_e = rtemp1/6./30.
        do is = 1, 4
            in = neighbor(is,ie)
            if (in > 0) then
                ! the mesher recomputes and applies the permutation four times
                do m = 1, 4
                do i = 1, NpF
                    if ((abs(vx(ibt(i,is,ie)) - vx(ibn(i,is,ie))) < epsil_e) .and. &
                        (abs(vy(ibt(i,is,ie)) - vy(ibn(i,is,ie))) < epsil_e) .and. &
                        (abs(vz(ibt(i,is,ie)) - vz(ibn(i,is,ie))) < epsil_e)) then
                        perm(i) = i
                    else
                        c = 1
                        do k = 1, NpF
                            if ((abs(vx(ibt(i,is,ie)) - vx(ibn(k,is,ie))) < epsil_e) .and. &
                                (abs(vy(ibt(i,is,ie)) - vy(ibn(k,is,ie))) < epsil_e) .and. &
                                (abs(vz(ibt(i,is,ie)) - vz(ibn(k,is,ie))) < epsil_e)) then
                                perm(i) = c
                            end if
                            c = c+1
                        end do
                    end if
                end do
                ibn(:,is,ie) = ibn(perm(:),is,ie)
                end do
            end if
        end do
    end do

    ! dtfactor (src/meshMod.f90:1134-1157)
    rtemp1 = 1e7
    do ie = 1, nelem
        min1 = minval(jac_all(:,ie))
        min2 = minval(sj_all(:,ie))
        if ((min1/min2) < rtemp1) rtemp1 = min1/min2
    end do
    rtemp3 = 1e7
    do ie = 1, nelem
        min1 = 2.0/3.0 * minGLL * (rtemp1/vpu(ie))
        if (min1 < rtemp3) rtemp3 = min1
    end do

    ! sources and receivers
    par%log = .false.
    src%nsrc = nsrc
    allocate(src%srcxyz(3,nsrc), src%srcrst(3,nsrc), src%srcelem(nsrc), src%srci(nsrc))
    src%srcxyz = real(src8, CUSTOM_REAL)
    src%srcrst = 0
    src%srcelem = 0
    src%srci = 0
    if (nsrc > 0) call findSource(par, src, vx, vy, vz, nglob, nelem, ibool, coord, elem, dr, ds, dt)
    rec%nrec = nrec
    allocate(rec%recxyz(3,nrec), rec%recrst(3,nrec), rec%recelem(nrec), rec%reci(nrec), rec%recnr(nrec))
    rec%recxyz = real(rec8, CUSTOM_REAL)
    rec%recrst = 0
    rec%recelem = 0
    rec%reci = 0
    do i = 1, nrec
        rec%recnr(i) = i
    end do
    if (nrec > 0) call findReceiver(par, rec, vx, vy, vz, nglob, nelem, ibool, coord, elem, dr, ds, dt)

    open(unit=32, file=trim(outfile), form='unformatted', access='stream', status='replace')
    write(32) int(Np, 4), int(NpF, 4), int(CUSTOM_REAL, 4), int(nMB, 4)
    write(32) int(neighbor, 4), int(face, 4)
    write(32) real(rtemp3, 8), real(minGLL*rtemp1, 8)
    write(32) int(src%srcelem, 4), real(src%srcrst, 8), int(rec%recelem, 4), real(rec%recrst, 8)
    do l = 1, nsel
        ie = sel(l)
        iv = ibool(:,ie)
        call geometricFactors3d(rx, sx, tx, ry, sy, ty, rz, sz, tz, jac, vx(iv), vy(iv), vz(iv), dr, ds, dt)
        call normals3d(nx, ny, nz, sj, dr, ds, dt, vx(iv), vy(iv), vz(iv), fmask)
        ! fscale (src/meshMod.f90:878-889)
        c = 1
        do k = 1, 4
            do j = 1, NpF
                d = fmask(j,k)
                fscale(c) = sj(c)/jac(d)
                c = c+1
            end do
        end do
        ! face frames (src/meshMod.f90:843-876)
        do is = 1, 4
            if (is == 1) then
                c = 1; d = 2; e = 0*NpF+1
            else if (is == 2) then
                c = 1; d = 2; e = 1*NpF+1
            else if (is == 3) then
                c = 2; d = 3; e = 2*NpF+1
            else
                c = 1; d = 3; e = 3*NpF+1
            end if
            nnormal(1,is) = nx(e)
            nnormal(2,is) = ny(e)
            nnormal(3,is) = nz(e)
            rtemp1 = coord(1,elem(d,ie)) - coord(1,elem(c,ie))
            rtemp2 = coord(2,elem(d,ie)) - coord(2,elem(c,ie))
            rtemp3 = coord(3,elem(d,ie)) - coord(3,elem(c,ie))
            rtemp4 = sqrt(rtemp1**2+rtemp2**2+rtemp3**2)
            stangential(1,is) = rtemp1/rtemp4
            stangential(2,is) = rtemp2/rtemp4
            stangential(3,is) = rtemp3/rtemp4
            rtemp1 = nnormal(2,is)*stangential(3,is) - nnormal(3,is)*stangential(2,is)
            rtemp2 = nnormal(3,is)*stangential(1,is) - nnormal(1,is)*stangential(3,is)
            rtemp3 = nnormal(1,is)*stangential(2,is) - nnormal(2,is)*stangential(1,is)
            rtemp4 = sqrt(rtemp1**2+rtemp2**2+rtemp3**2)
            ttangential(1,is) = rtemp1/rtemp4
            ttangential(2,is) = rtemp2/rtemp4
            ttangential(3,is) = rtemp3/rtemp4
        end do
        write(32) real(vx(iv), 8), real(vy(iv), 8), real(vz(iv), 8)
        write(32) real(rx, 8), real(ry, 8), real(rz, 8), real(sx, 8), real(sy, 8), real(sz, 8)
        write(32) real(tx, 8), real(ty, 8), real(tz, 8), real(jac, 8)
        write(32) real(nx, 8), real(ny, 8), real(nz, 8), real(sj, 8), real(fscale, 8)
        write(32) real(nnormal, 8), real(stangential, 8), real(ttangential, 8)
        write(32) int(ibt(:,:,ie), 4), int(ibn(:,:,ie), 4)
    end do
    f0_att = real(f0_8, CUSTOM_REAL)
    fr_att = real(fr_8, CUSTOM_REAL)
    call anelasticPreCalc(f0_att, fr_att, wl, wk)
    write(32) real(wl, 8), real(wk, 8)
    do m = 1, nmat
        qp = real(mat8(1,m), CUSTOM_REAL)
        qs = real(mat8(2,m), CUSTOM_REAL)
        vp = real(mat8(3,m), CUSTOM_REAL)
        vs = real(mat8(4,m), CUSTOM_REAL)
        rho = real(mat8(5,m), CUSTOM_REAL)
        call calcAnelasticCoefficients(qp, qs, vp, vs, rho, f0_att, fr_att, wl, wk, ylambda, ymu, lambda_u, mu_u)
        write(32) real(ylambda, 8), real(ymu, 8), real(lambda_u, 8), real(mu_u, 8)
    end do
    close(32)
end program ref_geom_probe
