! ref_glue_assembly.F90 -- TEST INFRASTRUCTURE ONLY.
! Glue compiled with the reference's own sources (oracle/Makefile) into
! oracle/_ref/libref_assembly.so.  It fills the reference's module state
! (modules variables and arrays, src2LOWNS/0mod-*.F90) from box arrays,
! then calls the reference routines unchanged, in the order a time step's
! first pressure-correction pass runs them:
!   set_nflag                     (src2LOWNS/nfd2015.F90:7)
!   init_matrix                   (src2LOWNS/1sub-matrix.F90:3)
!   set_initial_surface_position  (src2LOWNS/1sub-normal-ini.F90:6)
!   cal_err_of_continuity, ik = 0 (src2LOWNS/1sub-momentum.F90:235)
!   cal_matrix_element            (src2LOWNS/1sub-momentum.F90:465)
! and copies the resulting flags, row numbering and PCGPME inputs
! (alu, llu, de, b) back.  The cell numbering nn restates the loops of
! read_data (src2LOWNS/data_io.F90:168-199): solid value 0 (flow cells) and
! the boundary-object codes are numbered in one (j, k, i) sweep, the
! appended object codes (-1, -128) after it; solid value 1 (outside the
! case) gets mn = 0.  Object cells keep f = -1 and b from their code
! (b_of below).  dx and x are allocated
! with bounds that cover every index the routines touch (the reference's
! own allocation, 0mod-arrays.F90:205-206, starts at 1 while set_dx writes
! from ls-2).
!
! Box convention (shared with the C oracle and the product): cells
! i=0..nx+1, j=0..ny+1, k=0..nz+1 with the ring solid; box index
! c = i + X*(j + Y*k), X = nx+2, Y = ny+2.
subroutine ref_build_system(nx, ny, nz, solid, fvv, fbv, av, unv, dxv, xfv, stride, dt_in, &
                            nf_out, nb_out, row2cell, lda, AL, LLo, Dd, Bb, nrows, mp_out) &
                            bind(C, name="ref_build_system")
    use iso_c_binding
    use omp_lib, only: omp_set_num_threads
    use variables, only: is, js, ks, ie, je, ke, inns, inne, inne2, dt, mnoe, m_alloc, ik, &
                         nl, mp, mms, NTHR, icon
    use arrays, only: ls, le, lf, dim3d, dx, x, mn, nff, in, jn, kn, u, un, f, fp, p, ox, sfn, &
                      fln, a, fb, a0, fb0, sp, spn, nor, objno, alu_c, ll_c, de_c, b, de, xe, &
                      alu, mno, llu, mno2mn, alloc_uvwfp, alloc_afb, alloc_a0fb0, alloc_norsp
    implicit none
    integer(c_int), value :: nx, ny, nz, stride, lda
    integer(c_int8_t), intent(in) :: solid(0:*)
    real(c_double), intent(in) :: fvv(0:*), fbv(0:*), av(0:*), unv(0:*), dxv(0:*), xfv(0:*)
    real(c_double), value :: dt_in
    integer(c_int8_t), intent(out) :: nf_out(0:*), nb_out(0:*)
    integer(c_int), intent(out) :: row2cell(0:*), LLo(lda, 6), nrows
    real(c_double), intent(out) :: AL(lda, 6), Dd(lda), Bb(lda), mp_out
    interface
        subroutine set_nflag()
        end subroutine
        subroutine init_matrix()
        end subroutine
        subroutine set_initial_surface_position()
        end subroutine
        subroutine cal_err_of_continuity()
        end subroutine
        subroutine cal_matrix_element()
        end subroutine
    end interface
    integer :: XX, YY, ncell, i, j, k, c, nn, l, r, mx2

    XX = nx + 2 ; YY = ny + 2 ; ncell = XX * YY * (nz + 2)
    ! One OpenMP thread: set_initial_surface_position's parallel loop
    ! (1sub-normal-ini.F90:17) shares its temporaries fff, xx, dxx between
    ! threads, and set_nflag's loops are order dependent (nfd2015.F90:133-153);
    ! the serial result is the defined one.
    NTHR = 1
    call omp_set_num_threads(1)
    is = 1 ; js = 1 ; ks = 1 ; ie = nx + 1 ; je = ny + 1 ; ke = nz + 1
    ls = [is, js, ks] ; le = [ie, je, ke]
    lf(0,0:2) = [1,0,0] ; lf(1,0:2) = [0,1,0] ; lf(2,0:2) = [0,0,1]
    dim3d = .true.
    icon = 0

    ! fresh module state
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
    if (allocated(alu_c)) deallocate(alu_c, ll_c, de_c)
    if (allocated(b)) deallocate(b, de, xe, alu, mno, llu, mno2mn)
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

    ! cell numbering nn (data_io.F90:139-158): j, k, i order, flow cells only
    allocate(mn(is-2:ie+2, ks-2:ke+2, js-2:je+2)) ; mn = 0
    allocate(in(ncell), jn(ncell), kn(ncell))
    ! numbering as read_data's (data_io.F90:168-199): flow cells and boundary
    ! objects in one (j, k, i) sweep, then the appended object cells
    nn = 0
    do j = js-1, je
        do k = ks-1, ke
            do i = is-1, ie
                c = i + XX*(j + YY*k)
                if (solid(c) == 0 .or. in_main(solid(c))) then
                    nn = nn + 1 ; mn(i,k,j) = nn ; in(nn) = i ; jn(nn) = j ; kn(nn) = k
                endif
            enddo
        enddo
    enddo
    inns = 1 ; inne = nn
    do j = js-1, je
        do k = ks-1, ke
            do i = is-1, ie
                c = i + XX*(j + YY*k)
                if (solid(c) == -1 .or. solid(c) == -128) then
                    nn = nn + 1 ; mn(i,k,j) = nn ; in(nn) = i ; jn(nn) = j ; kn(nn) = k
                endif
            enddo
        enddo
    enddo
    inne2 = nn
    allocate(nff(0:inne2+1))
    nff(:)%f = 0 ; nff(:)%b = 0 ; nff(:)%fp = 0 ; nff(:)%bp = 0
    nff(0)%f = -1 ; nff(0)%b = 0 ; nff(inne2+1)%f = -1 ; nff(inne2+1)%b = -1
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        if (solid(c) /= 0) then
            nff(nn)%f = -1 ; nff(nn)%b = b_of(solid(c))
        endif
    enddo
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

    call set_nflag
    call init_matrix
    call set_initial_surface_position
    dt = dt_in
    ik = 0
    call cal_err_of_continuity
    call cal_matrix_element

    mp_out = mp
    nrows = mnoe
    nf_out(0:ncell-1) = -1 ; nb_out(0:ncell-1) = 0
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        nf_out(c) = int(nff(nn)%f, c_int8_t)
        if (nff(nn)%f == -1) then
            nb_out(c) = code_of(nff(nn)%b, nn > inne)
        else
            nb_out(c) = int(nff(nn)%b, c_int8_t)
        endif
    enddo
    do r = 1, mnoe
        nn = mno2mn(r)
        row2cell(r-1) = in(nn) + XX*(jn(nn) + YY*kn(nn))
        AL(r, 1:6) = alu(r, 1:6)
        LLo(r, 1:6) = llu(r, 1:6)
        Dd(r) = de(r)
        Bb(r) = b(r)
    enddo
contains
    ! object codes (the box grids' solid values other than 0 and 1): a
    ! boundary object with nfb = b (1 <= b < 200) has code b for 2..126, -2
    ! for b = 1 and 124 - b (-3..-75) for b >= 127; the cells appended with
    ! nfb 300 / 200 have -1 / -128 (their b becomes -1 / 0)
    logical function in_main(s)
        integer(c_int8_t), intent(in) :: s
        in_main = s /= 0 .and. s /= 1 .and. s /= -1 .and. s /= -128
    end function
    integer function b_of(s)
        integer(c_int8_t), intent(in) :: s
        if (s == -128) then
            b_of = 0
        elseif (s == -1) then
            b_of = -1
        elseif (s == -2) then
            b_of = 1
        elseif (s < 0) then
            b_of = 124 - s
        else
            b_of = s
        endif
    end function
    integer(c_int8_t) function code_of(b, appended)
        integer, intent(in) :: b
        logical, intent(in) :: appended
        if (appended) then
            if (b == -1) then
                code_of = -1
            else
                code_of = -128
            endif
        elseif (b == 1) then
            code_of = -2
        elseif (b >= 127) then
            code_of = int(124 - b, c_int8_t)
        else
            code_of = int(b, c_int8_t)
        endif
    end function
end subroutine ref_build_system

! One pressure-correction pass of the reference after ref_build_system on
! the same grid (module state persists): xe(1:mnoe) = the given solution,
! xe(0) = 0, then p(mno2mn(r)) -= xe(r) (1sub-momentum.F90:53-58),
! estimate_velocity2 (1sub-momentum.F90:590-651), and the next pass's
! cal_err_of_continuity with ik > 0 (RHS and DMAX for the same numbering).
! Returns un (box layout, 3*ncell), p (per cell) and b, DMAX.
subroutine ref_correction_pass(nx, ny, nz, xe_in, unv, pv, Bb, dmax_out) &
        bind(C, name="ref_correction_pass")
    use iso_c_binding
    use variables, only: inne2, mnoe, ik, dmax
    use arrays, only: in, jn, kn, un, xe, b, p, mno2mn
    implicit none
    integer(c_int), value :: nx, ny, nz
    real(c_double), intent(in) :: xe_in(*)
    real(c_double), intent(inout) :: unv(0:*), pv(0:*)
    real(c_double), intent(out) :: Bb(*), dmax_out
    interface
        subroutine estimate_velocity2()
        end subroutine
        subroutine cal_err_of_continuity()
        end subroutine
    end interface
    integer :: XX, YY, ncell, nn, l, c, r
    XX = nx + 2 ; YY = ny + 2 ; ncell = XX * YY * (nz + 2)
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        p(nn) = pv(c)
    enddo
    xe(0) = 0d0
    xe(1:mnoe) = xe_in(1:mnoe)
    do r = 1, mnoe
        p(mno2mn(r)) = p(mno2mn(r)) - xe(r)
    enddo
    call estimate_velocity2
    ik = 1
    call cal_err_of_continuity
    do nn = 1, inne2
        c = in(nn) + XX*(jn(nn) + YY*kn(nn))
        pv(c) = p(nn)
        do l = 0, 2
            unv(l*ncell + c) = un(l, nn)
        enddo
    enddo
    Bb(1:mnoe) = b(1:mnoe)
    dmax_out = dmax
end subroutine ref_correction_pass

! set_dx (src2LOWNS/1sub-main.F90:53-74) on one axis: x(ls-2 .. le+1) in,
! dx(ls-2 .. le+1) out, with ls = 1 and le = m - 3 on axis 0.  The module
! arrays are allocated over that whole range (the reference's own
! allocation, 0mod-arrays.F90:205-206, starts at 1).
subroutine ref_set_dx(m, xv, dxv) bind(C, name="ref_set_dx")
    use iso_c_binding
    use arrays, only: ls, le, dx, x
    implicit none
    integer(c_int), value :: m
    real(c_double), intent(inout) :: xv(0:m-1)
    real(c_double), intent(out) :: dxv(0:m-1)
    interface
        subroutine set_dx()
        end subroutine
    end interface
    integer :: i
    if (allocated(dx)) deallocate(dx)
    if (allocated(x)) deallocate(x)
    allocate(dx(0:2, -1:m-2), x(0:2, -1:m-2))
    dx = 0d0
    x = 0d0
    ls = [1, 1, 1]
    le = [m - 3, 2, 2]
    do i = 0, m - 1
        x(0, i - 1) = xv(i)
    enddo
    ! axes 1 and 2: a small uniform grid so their loops stay in bounds
    do i = -1, 3
        x(1, i) = dble(i)
        x(2, i) = dble(i)
    enddo
    call set_dx()
    do i = 0, m - 1
        dxv(i) = dx(0, i - 1)
        xv(i) = x(0, i - 1)
    enddo
end subroutine ref_set_dx
