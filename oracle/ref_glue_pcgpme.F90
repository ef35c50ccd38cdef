! ref_glue_pcgpme.F90 -- TEST INFRASTRUCTURE ONLY.
! Glue compiled together with the reference's own src2LOWNS/pcgpme_para.F90
! and its modules into oracle/_ref/libref_pcgpme.so.  PCGPME itself is an
! external subroutine (pcgpme_para.F90:1), so C calls it directly as pcgpme_;
! this file only sets the module thread count NTHR (0mod-variables.F90:83)
! that PCGPME's OpenMP schedules divide by.
subroutine ref_set_threads(n) bind(C, name="ref_set_threads")
    use iso_c_binding, only: c_int
    use variables, only: NTHR
    use omp_lib, only: omp_set_num_threads
    implicit none
    integer(c_int), value :: n
    NTHR = n
    call omp_set_num_threads(n)
end subroutine ref_set_threads

! Flush unit 6: the reference writes progress lines there (e.g. 'bn0=' on
! IER = 3000, pcgpme_para.F90:250); a caller that silences stdout around a
! call flushes them while silenced.
subroutine ref_flush() bind(C, name="ref_flush")
    flush(6)
end subroutine ref_flush
