! pcgpme_will_shim_test.F90 -- a Fortran caller of the PCGPME_WILL drop-in,
! as disp_adjust calls it (src2LOWNS/1sub-2DH-momentum.f90:325-326):
! NL = 4, NA = N, N2 = 64, EPS = 1e-10, PARM = 5, ITER = 1000.  The operator
! is the 2DH linear-dispersion matrix built as disp_param builds it
! (src2LOWNS/1sub-2DH-initial_reset.f90:472-505, A = 0.4): one row per cell
! with ZK < 0 (numbered i fastest), x neighbours skipped on the first and
! last column, y neighbours on the first and last row, packed LL with zeros
! after the last neighbour, DIAG = -1 - sum(AL).  Non-uniform spacing, a
! sloping bed with dry land at large i; PSI = 0 on rows without neighbours
! as in disp_adjust.  Prints N, the max-norm of D*B + AL*B(LL) - PSI for the
! returned B, and whether AL, LL, D changed.
program pcgpme_will_shim_test
    implicit none
    integer, parameter :: nx = 48, ny = 36
    real(8), parameter :: A = 0.4d0
    integer :: mnd(0:nx+1,0:ny+1), N, NL, NA, N2, ITER, i, j, r, jj, ii, nm, np, k, changed
    real(8) :: ZK(0:nx+1,0:ny+1), DX(2,0:nx+1,0:ny+1), EPS, PARM, res, s, bmax
    real(8), allocatable :: AL(:,:), D(:), B(:), A0(:,:), D0(:), PSI(:)
    integer, allocatable :: LL(:,:), L0(:,:)
    do j = 0, ny+1
    do i = 0, nx+1
        ZK(i,j) = -(6d0 - 0.18d0*i) + 0.4d0*sin(0.3d0*j)
        DX(1,i,j) = 1d0 + 0.1d0*sin(0.7d0*i)
        DX(2,i,j) = 0.8d0 + 0.05d0*cos(0.5d0*j)
    enddo
    enddo
    mnd = 0 ; N = 0
    do j = 1, ny
    do i = 1, nx
        if (ZK(i,j) < 0d0) then
            N = N + 1 ; mnd(i,j) = N
        endif
    enddo
    enddo
    NL = 4 ; NA = N ; N2 = 64
    allocate(AL(NA,NL), LL(NA,NL), D(N), B(N), A0(NA,NL), L0(NA,NL), D0(N), PSI(N))
    AL = 0d0 ; LL = 0 ; D = -1d0
    do j = 1, ny
    do i = 1, nx
        r = mnd(i,j)
        if (r == 0) cycle
        jj = 0
        do ii = 1, 2
            if (ii == 1) then
                if (i == 1 .or. i == nx) cycle
                nm = mnd(i-1,j) ; np = mnd(i+1,j)
            else
                if (j == 1 .or. j == ny) cycle
                nm = mnd(i,j-1) ; np = mnd(i,j+1)
            endif
            if (nm /= 0) then
                jj = jj + 1 ; LL(r,jj) = nm
                if (ii == 1) then
                    AL(r,jj) = A*2d0*ZK(i,j)*ZK(i,j)/DX(1,i,j)/(DX(1,i,j)+DX(1,i-1,j))
                else
                    AL(r,jj) = A*2d0*ZK(i,j)*ZK(i,j)/DX(2,i,j)/(DX(2,i,j)+DX(2,i,j-1))
                endif
            endif
            if (np /= 0) then
                jj = jj + 1 ; LL(r,jj) = np
                if (ii == 1) then
                    AL(r,jj) = A*2d0*ZK(i,j)*ZK(i,j)/DX(1,i,j)/(DX(1,i,j)+DX(1,i+1,j))
                else
                    AL(r,jj) = A*2d0*ZK(i,j)*ZK(i,j)/DX(2,i,j)/(DX(2,i,j)+DX(2,i,j+1))
                endif
            endif
        enddo
        D(r) = D(r) - sum(AL(r,:))
        ! a row without neighbours gets no hxx term in disp_adjust
        ! (1sub-2DH-momentum.f90:254-306), so PSI = 0 there; PCGPME_WILL
        ! returns such a row's final residual, not its X (pcgpme_2DH.f90:258-266)
        PSI(r) = 0d0
        if (jj > 0) PSI(r) = 1d-3*sin(0.21d0*i)*cos(0.17d0*j)
    enddo
    enddo
    A0 = AL ; L0 = LL ; D0 = D ; B = PSI
    EPS = 1d-10 ; PARM = 5d0 ; ITER = 1000
    call PCGPME_WILL(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER)
    res = 0d0 ; bmax = 0d0
    do r = 1, N
        s = D0(r)*B(r) - PSI(r)
        do k = 1, NL
            if (L0(r,k) > 0) s = s + A0(r,k)*B(L0(r,k))
        enddo
        res = max(res, abs(s))
        bmax = max(bmax, abs(B(r)))
    enddo
    changed = 0
    if (any(AL /= A0) .or. any(LL /= L0) .or. any(D /= D0)) changed = 1
    write(*,'(A,I8,A,I2,A,ES12.4,A,ES12.4,A,ES12.4)') 'N=', N, ' CHANGED=', changed, &
        ' RES=', res, ' BMAX=', bmax, ' PSIMAX=', maxval(abs(PSI))
end program pcgpme_will_shim_test
