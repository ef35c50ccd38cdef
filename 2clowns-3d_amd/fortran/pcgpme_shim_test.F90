! pcgpme_shim_test.F90 -- a Fortran caller of the drop-in, as
! cal_poisson_equation (src2LOWNS/1sub-momentum.F90:553-588) calls PCGPME:
! iter=10000, eps=(1e-10,0), parm=5, N2=64, NA=N+60, X zeroed.  The operator
! is the uniform-grid pressure matrix (dx=dy=1, dz=0.5, mp=0.25) on an
! nx*ny*nz box with closed walls and a top layer of free-surface rows.
! Prints N, ITER, IER, EPS and the max-norm of the unscaled residual
! b - A x, then the scaled coefficients of one interior row.
program pcgpme_shim_test
    implicit none
    integer, parameter :: nx = 24, ny = 20, nz = 12
    integer :: N, NA, NL, N2, ITER, IER, i, j, k, r, nb, slot
    real(8), allocatable :: AL(:,:), D(:), B(:), X(:), A0(:,:), D0(:), B0(:)
    integer, allocatable :: LL(:,:), L0(:,:)
    real(8) :: EPS(2), PARM, res, s, f
    N = nx*ny*nz ; NA = N + 60 ; NL = 6 ; N2 = 64
    allocate(AL(NA,NL), LL(NA,NL), D(N), B(N), X(N+N2))
    allocate(A0(NA,NL), L0(NA,NL), D0(N), B0(N))
    AL = 0d0 ; LL = 0 ; D = 0d0
    do k = 1, nz
    do j = 1, ny
    do i = 1, nx
        r = i + nx*((j-1) + ny*(k-1))
        if (k == nz) then
            D(r) = 1d0
            f = 0.75d0
            LL(r,5) = r - nx*ny
            AL(r,5) = -(f-0.5d0)*0.5d0/(f*0.5d0+0.5d0*0.5d0)
            if (i > 1)  then ; LL(r,1) = r-1     ; AL(r,1) = -1d-20 ; endif
            if (i < nx) then ; LL(r,2) = r+1     ; AL(r,2) = -1d-20 ; endif
            if (j > 1)  then ; LL(r,3) = r-nx    ; AL(r,3) = -1d-20 ; endif
            if (j < ny) then ; LL(r,4) = r+nx    ; AL(r,4) = -1d-20 ; endif
        else
            do slot = 1, 6
                nb = 0
                select case (slot)
                case (1) ; if (i > 1)  nb = r-1
                case (2) ; if (i < nx) nb = r+1
                case (3) ; if (j > 1)  nb = r-nx
                case (4) ; if (j < ny) nb = r+nx
                case (5) ; if (k > 1)  nb = r-nx*ny
                case (6) ; nb = r+nx*ny
                end select
                if (nb > 0) then
                    LL(r,slot) = nb
                    if (slot <= 4) then
                        AL(r,slot) = -0.25d0*(2d0/1d0/(1d0+1d0))
                    else
                        AL(r,slot) = -0.25d0*(2d0/0.5d0/(0.5d0+0.5d0))
                    endif
                    D(r) = D(r) - AL(r,slot)
                endif
            enddo
        endif
        B(r) = sin(0.37d0*r) * merge(0d0, 1d0, k == nz)
    enddo
    enddo
    enddo
    A0 = AL ; L0 = LL ; D0 = D ; B0 = B
    X = 0d0 ; EPS(1) = 1d-10 ; EPS(2) = 0d0 ; PARM = 5d0 ; ITER = 10000
    call PCGPME(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER)
    res = 0d0
    do r = 1, N
        s = D0(r)*X(r) - B0(r)
        do slot = 1, NL
            if (L0(r,slot) > 0) s = s + A0(r,slot)*X(L0(r,slot))
        enddo
        res = max(res, abs(s))
    enddo
    r = (nx/2) + nx*((ny/2-1) + ny*(nz/2-1))
    write(*,'(A,I8,A,I6,A,I6,A,ES12.4,A,ES12.4)') 'N=', N, ' ITER=', ITER, ' IER=', IER, &
        ' EPS=', EPS(1), ' RES=', res
    write(*,'(A,7ES24.16)') 'ROW ', D(r), AL(r,1:6)
    write(*,'(A,6I10)') 'LL  ', LL(r,1:6)
end program pcgpme_shim_test
