! ref_main_decks.F90 -- TEST INFRASTRUCTURE ONLY.
! A main program around the deck glue (ref_glue_decks.F90), linked with the
! reference's own src2LOWNS/data_io.F90 into oracle/_ref/ref_decks_be and
! compiled with flang's -fconvert=big-endian, so that read_data, read_fdata
! and output_data (data_io.F90:87-364, 708-783) read and write big-endian
! records as the reference's own build does (src2LOWNS/Makefile:5).  flang's
! runtime takes the conversion only from a main program, which is why this
! is an executable and not part of libref_decks.so.
!
! Usage: ref_decks_be grid data fdata nmax xstride out [write_path t dt]
!   out: what ref_read_deck returns, one value a line in its argument order
!        (ext, counts, hdr, xo, then per numbered cell: i j k, U, F, P, a0,
!        fb0, nf nfb); reals as the 16 hexadecimal digits of their bits.
!   write_path: output_data then writes the data file of the state read, at
!        time t with step dt (each given as 16 hexadecimal digits).
program ref_decks_be
    use iso_c_binding
    implicit none
    interface
        subroutine ref_read_deck(gp, gl, dp, dl, fpath, fl, ext, counts, hdr, xo, xstride, nmax, cells, uo, &
                                 fo, po, ao, fbo, flags) bind(C, name="ref_read_deck")
            import :: c_char, c_int, c_double
            character(kind=c_char), intent(in) :: gp(*), dp(*), fpath(*)
            integer(c_int), value :: gl, dl, fl, xstride, nmax
            integer(c_int), intent(out) :: ext(6), counts(2), cells(3, *), flags(2, *)
            real(c_double), intent(out) :: hdr(4), xo(0:xstride-1, 0:2), uo(0:2, *), fo(*), po(*), &
                                           ao(0:2, *), fbo(*)
        end subroutine
        subroutine ref_write_data(tp, tl, t_in, dt_in) bind(C, name="ref_write_data")
            import :: c_char, c_int, c_double
            character(kind=c_char), intent(in) :: tp(*)
            integer(c_int), value :: tl
            real(c_double), value :: t_in, dt_in
        end subroutine
    end interface
    character(len=1024) :: gpath, dpath, fpath, opath, wpath, arg
    integer(c_int) :: nmax, xs, ext(6), counts(2)
    integer(c_int), allocatable :: cells(:, :), flags(:, :)
    real(c_double) :: hdr(4), t, dt
    real(c_double), allocatable :: xo(:, :), uo(:, :), fo(:), po(:), ao(:, :), fbo(:)
    integer(c_int64_t) :: bits
    integer :: i, l, nn, u

    call get_command_argument(1, gpath)
    call get_command_argument(2, dpath)
    call get_command_argument(3, fpath)
    call get_command_argument(4, arg)
    read (arg, *) nmax
    call get_command_argument(5, arg)
    read (arg, *) xs
    call get_command_argument(6, opath)
    allocate (cells(3, nmax), flags(2, nmax), xo(0:xs-1, 0:2), uo(0:2, nmax), fo(nmax), po(nmax), &
              ao(0:2, nmax), fbo(nmax))
    call ref_read_deck(gpath, len_trim(gpath), dpath, len_trim(dpath), fpath, len_trim(fpath), ext, counts, &
                       hdr, xo, xs, nmax, cells, uo, fo, po, ao, fbo, flags)

    open (newunit=u, file=trim(opath), form='formatted', status='replace')
    do i = 1, 6
        write (u, '(I0)') ext(i)
    enddo
    write (u, '(I0)') counts(1)
    write (u, '(I0)') counts(2)
    do i = 1, 4
        call put(hdr(i))
    enddo
    do l = 0, 2
        do i = 0, xs - 1
            call put(xo(i, l))
        enddo
    enddo
    do nn = 1, min(counts(2), nmax)
        write (u, '(I0)') cells(1, nn)
        write (u, '(I0)') cells(2, nn)
        write (u, '(I0)') cells(3, nn)
        call put(uo(0, nn))
        call put(uo(1, nn))
        call put(uo(2, nn))
        call put(fo(nn))
        call put(po(nn))
        call put(ao(0, nn))
        call put(ao(1, nn))
        call put(ao(2, nn))
        call put(fbo(nn))
        write (u, '(I0)') flags(1, nn)
        write (u, '(I0)') flags(2, nn)
    enddo
    close (u)

    if (command_argument_count() >= 9) then
        call get_command_argument(7, wpath)
        call get_command_argument(8, arg)
        read (arg, '(Z16)') bits
        t = transfer(bits, t)
        call get_command_argument(9, arg)
        read (arg, '(Z16)') bits
        dt = transfer(bits, dt)
        call ref_write_data(wpath, len_trim(wpath), t, dt)
    endif

contains

    subroutine put(v)
        real(c_double), intent(in) :: v
        write (u, '(Z16.16)') transfer(v, bits)
    end subroutine

end program ref_decks_be
