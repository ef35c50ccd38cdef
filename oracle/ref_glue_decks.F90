! ref_glue_decks.F90 -- TEST INFRASTRUCTURE ONLY.
! Glue compiled with the reference's own src2LOWNS/data_io.F90 and its
! modules (oracle/Makefile) into oracle/_ref/libref_decks.so.  It sets the
! module state the reference's case readers take (file names, drift off,
! one-dimensional coordinate arrays), calls read_data and read_fdata
! (data_io.F90:87-364) unchanged, and copies what they leave behind out:
! extents, coordinates, and per numbered cell nn its (i, j, k), U, F, P,
! a0/fb0 and flags.  ref_write_data calls output_data (data_io.F90:708-783)
! on that state.  The flang runtime reads and writes the host byte order
! here (no main program sets FORT_CONVERT), so the tests give it
! little-endian files.
subroutine ref_read_deck(gp, gl, dp, dl, fpath, fl, ext, counts, hdr, xo, xstride, nmax, cells, uo, fo, po, &
                         ao, fbo, flags) bind(C, name="ref_read_deck")
    use iso_c_binding
    use variables, only: inputxyz, input, INPUTf0, t0, g, nu, dt, is, js, ks, ie, je, ke, inne, inne2, &
                         NTHR, g_read, nu_read
    use arrays, only: ls, le, x, dx, mn, nff, in, jn, kn, nf3, nfb3, in2d, mn2d, u, un, f, fp, p, ox, &
                      sfn, fln, a0, fb0, a, fb, drift, one_dim_array, xyzn
    implicit none
    character(kind=c_char), intent(in) :: gp(*), dp(*), fpath(*)
    integer(c_int), value :: gl, dl, fl, xstride, nmax
    integer(c_int), intent(out) :: ext(6), counts(2), cells(3, *), flags(2, *)
    real(c_double), intent(out) :: hdr(4), xo(0:xstride-1, 0:2), uo(0:2, *), fo(*), po(*), ao(0:2, *), fbo(*)
    interface
        subroutine read_data()
        end subroutine
        subroutine read_fdata()
        end subroutine
    end interface
    integer :: i, l, nn
    NTHR = 1
    inputxyz = ' '
    input = ' '
    INPUTf0 = ' '
    do i = 1, gl
        inputxyz(i:i) = gp(i)
    enddo
    do i = 1, dl
        input(i:i) = dp(i)
    enddo
    do i = 1, fl
        INPUTf0(i:i) = fpath(i)
    enddo
    drift = .false.
    one_dim_array = .true.
    xyzn = .true.  ! module default; read_data clears it for a .xyz grid
    dt = 0d0
    g_read = 0d0
    nu_read = 0d0
    if (allocated(dx)) deallocate(dx)
    if (allocated(x)) deallocate(x)
    if (allocated(mn)) deallocate(mn)
    if (allocated(nff)) deallocate(nff)
    if (allocated(in)) deallocate(in, jn, kn)
    if (allocated(nf3)) deallocate(nf3)
    if (allocated(nfb3)) deallocate(nfb3)
    if (allocated(in2d)) deallocate(in2d)
    if (allocated(mn2d)) deallocate(mn2d)
    if (allocated(u)) deallocate(u, un, f, fp, p, ox, sfn, fln)
    if (allocated(a0)) deallocate(a0, fb0)
    if (allocated(a)) deallocate(a, fb)
    call read_data()
    call read_fdata()
    ext = [is, js, ks, ie, je, ke]
    counts = [inne, inne2]
    hdr = [t0, g, nu, dt]
    xo = 0d0
    do l = 0, 2
        do i = ls(l) - 2, le(l) + 1
            xo(i - (ls(l) - 2), l) = x(l, i)
        enddo
    enddo
    do nn = 1, min(inne2, nmax)
        cells(:, nn) = [in(nn), jn(nn), kn(nn)]
        uo(0:2, nn) = u(0:2, nn)
        fo(nn) = f(nn)
        po(nn) = p(nn)
        ao(0:2, nn) = a(0:2, nn)
        fbo(nn) = fb(nn)
        flags(:, nn) = [nff(nn)%f, nff(nn)%b]
    enddo
end subroutine ref_read_deck

subroutine ref_write_data(tp, tl, t_in, dt_in) bind(C, name="ref_write_data")
    use iso_c_binding
    use variables, only: title, t, dt, NTHR
    implicit none
    character(kind=c_char), intent(in) :: tp(*)
    integer(c_int), value :: tl
    real(c_double), value :: t_in, dt_in
    interface
        subroutine output_data()
        end subroutine
    end interface
    integer :: i
    NTHR = 1
    title = ' '
    do i = 1, tl
        title(i:i) = tp(i)
    enddo
    t = t_in
    dt = dt_in
    call output_data()
end subroutine ref_write_data
