! Chemistry driver for the reference's chem_module (test infrastructure).
!
! Compiled by oracle/Makefile (target chem) against
! plasma_pde/chemistry/fundconstants.f90 and a chemistry file of the
! reference (by path, never copied) into oracle/_ref/ref_chem_<name>; run only
! to write the fixtures under tests/golden/. It calls initializechemistry
! (the tables of assignreactions / setspecparams) and evaluates the module's
! functions on sample states.
!
! Input  (stream): int32 ns_states; real64 Tg, Pg; then per state real64 Te,
!        efield_mag, specden(nspecies).
! Output (stream): int32 nspecies, nreac, especnum, bgspecnum, ionspecmin,
!        ionspecmax; int32 reactants(nreac,nspecies), products(nreac,nspecies),
!        isratearrh(nreac) (0/1); real64 k_arrh(3,nreac), elecenergy(nreac),
!        molmass(nspecies), spec_charge(nspecies); then per state real64
!        rate(nreac) (getarrhrate or getcustomrate at Te), prod(nspecies)
!        (getspecproduction), mob(nspecies) and dco(nspecies) (getspecmobility
!        and getspecdcoeff for the charged species, 0 for the others), and for
!        the electrons dco of the neutral species that have one (0 otherwise),
!        inelterm (getelectroninelasticterm), collfreq
!        (getelectroncollisionfrequency).
program ref_chem
    use chem_module
    implicit none
    character(len=512) :: infile, outfile
    integer :: nst, s, r, j, isr(nreac)
    real*8 :: Tg, Pg, Te, ef, sd(nspecies), rate(nreac), prod(nspecies), mob(nspecies), dco(nspecies)
    real*8 :: inel, cf, p
    logical :: hasd(nspecies)

    call get_command_argument(1, infile)
    call get_command_argument(2, outfile)
    call initializechemistry()
    open(unit=21, file=trim(infile), access='stream', form='unformatted', status='old')
    open(unit=22, file=trim(outfile), access='stream', form='unformatted', status='replace')
    read(21) nst
    read(21) Tg, Pg
    do r = 1, nreac
        isr(r) = 0
        if (isratearrh(r)) isr(r) = 1
    enddo
    write(22) nspecies, nreac, especnum, bgspecnum, ionspecmin, ionspecmax
    write(22) reactants, products, isr
    write(22) k_arrh, elecenergy, molmass, spec_charge
    ! species with a diffusion coefficient: the charged ones and the
    ! solved neutrals
    do j = 1, nspecies
        hasd(j) = spec_charge(j) .ne. 0.d0 .or. (j .ge. neutralspecmin .and. j .le. neutralspecmax)
    enddo
    do s = 1, nst
        read(21) Te, ef, sd
        do r = 1, nreac
            if (isratearrh(r)) then
                rate(r) = getarrhrate(k_arrh(:, r), Te)
            else
                rate(r) = getcustomrate(r, k_arrh(:, r), Te, ef)
            endif
        enddo
        do j = 1, nspecies
            call getspecproduction(j, Te, Tg, sd, p, ef)
            prod(j) = p
            mob(j) = 0.d0
            dco(j) = 0.d0
            if (spec_charge(j) .ne. 0.d0) mob(j) = getspecmobility(j, sd, ef, Te, Tg, Pg)
            if (hasd(j)) dco(j) = getspecdcoeff(j, sd, ef, Te, Tg, Pg)
        enddo
        call getelectroninelasticterm(Te, Tg, sd, inel, ef)
        cf = getelectroncollisionfrequency(sd, ef, Te, Tg, Pg)
        write(22) rate, prod, mob, dco, inel, cf
    enddo
    close(21)
    close(22)
end program ref_chem
