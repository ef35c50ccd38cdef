! pcgpme_hip.F90 -- drop-in replacement for src2LOWNS/pcgpme_para.F90.
!
! Keeps the reference's external subroutine name and argument list,
!   subroutine PCGPME(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER)
! (pcgpme_para.F90:1,35-46), so cal_poisson_equation
! (src2LOWNS/1sub-momentum.F90:564) calls it unchanged.  The body forwards
! every argument by reference through an ISO_C_BINDING interface to
! clowns_pcgpme in libclowns_pcg.so (include/clowns_pcg.h), which runs the
! solve on the GPU and returns with X, B, AL, LL, D, EPS, ITER and IER
! updated as the reference leaves them.  Integers are default (4-byte)
! INTEGER as in the reference build (-r8 promotes only REAL).
subroutine PCGPME(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER)
    use iso_c_binding, only: c_double, c_int32_t
    implicit none
    integer, intent(in) :: NL, NA, N, N2
    real(8), dimension(NA,NL), intent(inout) :: AL
    integer, dimension(NA,NL), intent(inout) :: LL
    real(8), dimension(N), intent(inout) :: D, B
    real(8), dimension(N+N2), intent(inout) :: X
    real(8), dimension(2), intent(inout) :: EPS
    real(8), intent(in) :: PARM
    integer, intent(inout) :: ITER
    integer, intent(out) :: IER
    interface
        subroutine clowns_pcgpme(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER) &
                bind(C, name="clowns_pcgpme")
            import :: c_double, c_int32_t
            real(c_double) :: AL(*), D(*), B(*), X(*), EPS(*)
            integer(c_int32_t) :: NL, NA, N, N2, ITER, IER
            integer(c_int32_t) :: LL(*)
            real(c_double) :: PARM
        end subroutine clowns_pcgpme
    end interface
    call clowns_pcgpme(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER)
end subroutine PCGPME

! Replaces PCGPME_WILL (src2LOWNS/pcgpme_2DH.f90:1,28-41), the 2DH dispersion
! solve called by disp_adjust (1sub-2DH-momentum.f90:325-326), with the same
! name and argument list.  B returns what the reference leaves in it (the
! filtered solution, pcgpme_2DH.f90:255-268); AL, LL, D are not modified.
! The reference has no status argument: on non-convergence it writes a line
! to unit 6 (pcgpme_2DH.f90:248), as this shim does; a library error stops.
subroutine PCGPME_WILL(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER)
    use iso_c_binding, only: c_double, c_int32_t
    implicit none
    integer, intent(in) :: N, N2, NA, NL
    real(8), intent(in) :: PARM, EPS
    integer, intent(in) :: ITER
    real(8), dimension(NA,NL), intent(in) :: AL
    real(8), dimension(N), intent(in) :: D
    real(8), dimension(N), intent(inout) :: B
    integer, dimension(NA,NL), intent(in) :: LL
    integer :: IER
    interface
        subroutine clowns_pcgpme_will(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER,IER) &
                bind(C, name="clowns_pcgpme_will")
            import :: c_double, c_int32_t
            real(c_double) :: AL(*), D(*), B(*)
            integer(c_int32_t) :: NL, NA, N, N2, ITER, IER
            integer(c_int32_t) :: LL(*)
            real(c_double) :: EPS, PARM
        end subroutine clowns_pcgpme_will
    end interface
    call clowns_pcgpme_will(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER,IER)
    if (IER == 3000) then
        write(6,*) 'PCGPME_WILL: no convergence in ITER =', ITER
    else if (IER /= 0) then
        write(6,*) 'PCGPME_WILL: clowns_pcgpme_will failed, IER =', IER
        stop 1
    end if
end subroutine PCGPME_WILL
