! Test infrastructure only: never linked into the product library.
!
! Probe program linked against the reference's own Fortran modules (built by
! oracle/build_ref.sh from /root/reference/src).  It calls the reference's
! reference-element routines in the same sequence as the mesher does
! (src/meshMod.f90:601-663) and the solver's interpolation set-up
! (src/timeloopMod.f90:284-299) and dumps the results as raw little-endian
! arrays, so that the host-side precompute of this project can be compared
! against the reference operators bit for bit.
!
! Usage: ref_probe <outfile> [r s t]
program ref_probe
    use constantsMod
    use nodesMod
    use tetTrafoMod
    use vandermondeMod
    use dmatricesMod
    use liftMod
    use matrixMod
    use rosettaGammaMod
    implicit none
    real(kind=CUSTOM_REAL), dimension(Np) :: x, y, z, r, s, t, psi, w
    real(kind=CUSTOM_REAL), dimension(Np,Np) :: vdm, invvdm, mass, dr, ds, dt, vtinv
    real(kind=CUSTOM_REAL), dimension(Np,4*NpF) :: lift
    integer, dimension(NpF,4) :: fmask
    real(kind=CUSTOM_REAL), dimension(1) :: rr, ss, tt
    real(kind=CUSTOM_REAL) :: gam(16)
    character(len=256) :: outfile, arg
    integer :: i, c, d, nargs

    nargs = command_argument_count()
    call get_command_argument(1, outfile)
    rr(1) = -0.3; ss(1) = -0.2; tt(1) = -0.4
    if (nargs >= 4) then
        call get_command_argument(2, arg); read(arg,*) rr(1)
        call get_command_argument(3, arg); read(arg,*) ss(1)
        call get_command_argument(4, arg); read(arg,*) tt(1)
    end if

    call nodes3D(balpha(order), x, y, z)
    call xyzToRst(x, y, z, r, s, t)
    call vdm3d(vdm, r, s, t)
    call invVdm3D(vdm, invvdm, 0)
    mass = matmul(transpose(invvdm), invvdm)
    call dmatrices3d(dr, ds, dt, r, s, t, vdm)
    do i = 1, 4
        fmask(:, i) = 0
    end do
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
    call lift3d(lift, fmask, vdm, r, s, t)

    vtinv = transpose(vdm)
    call invert(vtinv)
    call vdm3D(psi, rr, ss, tt)
    w = matmul(vtinv, psi)

    do c = 1, 16
        gam(c) = lacz_gamma(real(c) * 0.5 + 0.5)
    end do

    open(unit=31, file=trim(outfile), form='unformatted', access='stream', status='replace')
    write(31) int(Np, 4), int(NpF, 4), int(CUSTOM_REAL, 4)
    write(31) x, y, z, r, s, t
    write(31) vdm, invvdm, mass, dr, ds, dt, lift
    write(31) fmask
    write(31) psi, w
    write(31) gam
    close(31)
end program ref_probe
