! Golden-vector driver for the reference's only Scharfetter-Gummel flux,
! find_sg_exponential_flux (mps1d/source/convdiff.f90:10-64).
!
! Test infrastructure: compiled here by oracle/Makefile against the mps1d
! sources (by path) into oracle/_ref/ref_sg1d; used only to write
! tests/golden/ fixtures.
!
! Input  (stream): int32 n, then n records of real64 (cL,cR,DL,DR,uL,uR,dx).
! Output (stream): n records of real64 (wL,wR,flux).
program ref_sg1d
    use convdiff
    implicit none
    character(len=512) :: infile, outfile
    integer :: n, i
    real*8 :: rec(7), wl, wr, flux

    call get_command_argument(1, infile)
    call get_command_argument(2, outfile)
    open(unit=21, file=trim(infile), access='stream', form='unformatted', status='old')
    open(unit=22, file=trim(outfile), access='stream', form='unformatted', status='replace')
    read(21) n
    do i = 1, n
        read(21) rec
        call find_sg_exponential_flux(rec(1), rec(2), rec(3), rec(4), rec(5), rec(6), rec(7), wl, wr, flux)
        write(22) wl, wr, flux
    enddo
    close(21)
    close(22)
end program ref_sg1d
