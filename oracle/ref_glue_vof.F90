! ref_glue_vof.F90 -- TEST INFRASTRUCTURE ONLY.
! Glue compiled with the reference's own sources (oracle/Makefile) into
! oracle/_ref/libref_vof.so.  Module state is filled from box arrays as in
! ref_glue_assembly.F90 (same numbering, same dx/x bounds), set_nflag
! (src2LOWNS/nfd2015.F90:7) classifies, and then the parts of cal_vof
! (src2LOWNS/1sub-vof.F90:10-38) run unchanged and in its order:
!   cal_quvwf, then quvwf_kaidan (mode 1; 1sub-vof.F90 is compiled without
!   NORMAL for that routine to exist) or the given PLIC face fluxes in place
!   of quvwf_normal (mode 0: qf(l,nn) = qs where qs_mask /= 0),
!   quvwf_modify, cal_new_f, simple_slide(0).
! Outputs: the flags and F after set_nflag (the VOF update's inputs), then
! qf, F and dt after the update, in the box layout, and the wall time of
! the update (cal_quvwf .. simple_slide) run at nthreads OpenMP threads.
subroutine ref_vof(nx, ny, nz, solid, fvv, fpv, fbv, av, unv, dxv, xfv, stride, dt_io, mode, qs, qs_mask, &
                   nf_out, nb_out, f_cls, qf_out, f_out, nthreads, secs) bind(C, name="ref_vof")
    use iso_c_binding
    use omp_lib, only: omp_set_num_threads, omp_get_wtime
    use variables, only: is, js, ks, ie, je, ke, inns, inne, inne2, dt, NTHR, icon, inn2d, n, t, m_alloc
    use arrays, only: ls, le, lf, dim3d, dx, x, mn, nff, in, jn, kn, u, un, f, fp, p, ox, sfn, &
                      fln, a, fb, a0, fb0, sp, spn, nor, objno, qf, in2d, alloc_uvwfp, alloc_afb, &
                      alloc_a0fb0, alloc_norsp, alloc_quvwf
    implicit none
    integer(c_int), value :: nx, ny, nz, stride, mode, nthreads
    integer(c_int8_t), intent(in) :: solid(0:*), qs_mask(0:*)
    real(c_double), intent(in) :: fvv(0:*), fpv(0:*), fbv(0:*), av(0:*), unv(0:*), dxv(0:*), xfv(0:*), qs(0:*)
    real(c_double), intent(inout) :: dt_io
    integer(c_int8_t), intent(out) :: nf_out(0:*), nb_out(0:*)
    real(c_double), intent(out) :: f_cls(0:*), qf_out(0:*), f_out(0:*), secs
    interface
        subroutine set_nflag()
        end subroutine
        subroutine cal_quvwf()
        end subroutine
        subroutine quvwf_kaidan()
        end subroutine
        subroutine quvwf_modify()
        end subroutine
        subroutine cal_new_f()
        end subroutine
        subroutine simple_slide(flag)
            integer, intent(in) :: flag
        end subroutine
    end interface
    integer :: XX, YY, ncell, i, j, k, c, nn, l, mx2
    real(c_double) :: t0

    XX = nx + 2 ; YY = ny + 2 ; ncell = XX * YY * (nz + 2)
    NTHR = 1
    call omp_set_num_threads(1)
    is = 1 ; js = 1 ; ks = 1 ; ie = nx + 1 ; je = ny + 1 ; ke = nz + 1
    ls = [is, js, ks] ; le = [ie, je, ke]
    lf(0,0:2) = [1,0,0] ; lf(1,0:2) = [0,1,0] ; lf(2,0:2) = [0,0,1]
    dim3d = .true.
    icon = 0 ; n = 1 ; t = 0d0

    if (allocated(dx)) deallocate(dx)
    if (allocated(x)) deallocate(x)
    if (allocated(mn)) deallocate(mn)
    if (allocated(nff)) deallocate(nff)
    if (allocated(in)) deallocate(in, jn, kn)
    if (allocated(u)) deallocate(u, un, f, fp, p, ox, sfn, fln)
    if (allocated(a)) deallocate(a, fb)
    if (allocated(a0)) deallocate(a0, fb0)
    if (allocated(sp)) deallocate(sp, spn, nor)
    if (allocated(objno)) deallocate(objno)
    if (allocated(qf)) deallocate(qf)
    if (allocated(in2d)) deallocate(in2d)
    m_alloc = 0

    mx2 = max(ie, je, ke) + 2
    allocate(dx(0:2, -2:mx2), x(0:2, -2:mx2))
    dx = 0d0 ; x = 0d0
    do l = 0, 2
        do i = 0, le(l)
            dx(l, i) = dxv(l*stride + i)
            x(l, i) = xfv(l*stride + i)
        enddo
        dx(l, -1) = dx(l, 0) ; x(l, -1) = x(l, 0) - dx(l, 0)
        dx(l, le(l)+1) = dx(l, le(l)) ; x(l, le(l)+1) = x(l, le(l)) + dx(l, le(l))
    enddo

    ! cell numbering nn (data_io.F90:160-176): j, k, i order, flow cells only
    allocate(mn(is-2:ie+2, ks-2:ke+2, js-2:je+2)) ; mn = 0
    allocate(in(ncell), jn(ncell), kn(ncell))
    nn = 0
    do j = js-1, je
        do k = ks-1, ke
            do i = is-1, ie
                c = i + XX*(j + YY*k)
                if (solid(c) == 0) then
                    nn = nn + 1 ; mn(i,k,j) = nn ; in(nn) = i ; jn(nn) = j ; kn(nn) = k
                endif
            enddo
        enddo
    enddo
    inns = 1 ; inne = nn ; inne2 = nn
    allocate(nff(0:inne2+1))
    nff(:)%f = 0 ; nff(:)%b = 0 ; nff(:)%fp = 0 ; nff(:)%bp = 0
    nff(0)%f = -1 ; nff(0)%b = 0 ; nff(inne2+1)%f = -1 ; nff(inne2+1)%b = -1
    call alloc_uvwfp(inne2)
    call alloc_afb(inne2, inne2)
    call alloc_a0fb0(inne2, inne2)
    call alloc_norsp(inne2)
    allocate(objno(0:inne2)) ; objno = 0
    u = 0d0 ; un = 0d0 ; f = 0d0 ; a = 0d0 ; fb = 0d0
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        f(nn) = fvv(c) ; fb(nn) = fbv(c) ; fb0(nn) = fbv(c)
        do l = 0, 2
            a(l, nn) = av(l*ncell + c)
            a0(l, nn) = av(l*ncell + c)
            un(l, nn) = unv(l*ncell + c)
        enddo
    enddo
    u = un
    ! columns of simple_slide (read_data, data_io.F90:178-183)
    allocate(in2d(0:1, (ie-is)*(je-js))) ; in2d = 0 ; inn2d = 0
    do j = js, je-1
        do i = is, ie-1
            inn2d = inn2d + 1 ; in2d(0:1, inn2d) = [i, j]
        enddo
    enddo

    call set_nflag
    nf_out(0:ncell-1) = -1 ; nb_out(0:ncell-1) = 0 ; f_cls(0:ncell-1) = 0d0
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        nf_out(c) = int(nff(nn)%f, c_int8_t)
        nb_out(c) = int(nff(nn)%b, c_int8_t)
        f_cls(c) = f(nn)
        fp(nn) = fpv(c)
    enddo

    dt = dt_io
    call alloc_quvwf(inne2)
    qf = 0d0
    ! the update itself at nthreads OpenMP threads (set_nflag above ran at one)
    call omp_set_num_threads(max(1, nthreads))
    t0 = omp_get_wtime()
    call cal_quvwf
    if (mode == 1) then
        call quvwf_kaidan
    else
        do nn = 1, inne
            c = in(nn) + XX*(jn(nn) + YY*kn(nn))
            do l = 0, 2
                if (qs_mask(l*ncell + c) /= 0) qf(l, nn) = qs(l*ncell + c)
            enddo
        enddo
    endif
    call quvwf_modify
    call cal_new_f
    call simple_slide(0)
    secs = omp_get_wtime() - t0
    call omp_set_num_threads(1)

    dt_io = dt
    qf_out(0:3*ncell-1) = 0d0
    f_out(0:ncell-1) = fvv(0:ncell-1)
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        f_out(c) = f(nn)
        do l = 0, 2
            qf_out(l*ncell + c) = qf(l, nn)
        enddo
    enddo
end subroutine ref_vof
