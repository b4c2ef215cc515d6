! Chemistry driver for the reference's chem_modules with the four-argument
! transport interface, getspecmobility(specnum,Te,Tg,Pg) and
! getspecdcoeff(specnum,Te,Tg,Pg) (plasma_pde/chemistry/H2_chem.f90 and
! O2_chem.f90; the plasma solver itself calls the six-argument interface of
! H2_O2_chem.f90 and ion_elec_chem.f90 and links with neither). Test
! infrastructure: compiled by oracle/Makefile (target chem4) against
! mgridlib/fundconstants.f90 (the constant names these modules use:
! mass_prot, mass_elec, echarge) and the chemistry file by path into
! oracle/_ref/ref_chem_<name>, and run only to write fixtures under
! tests/golden/. Same streams as ref_chem.f90: it calls initializechemistry
! and evaluates the module's functions on sample states.
!
! Input  (stream): int32 ns_states; real64 Tg, Pg; then per state real64 Te,
!        efield_mag, specden(nspecies).
! Output (stream): int32 nspecies, nreac, especnum, bgspecnum, ionspecmin,
!        ionspecmax; int32 reactants(nreac,nspecies), products(nreac,nspecies),
!        isratearrh(nreac) (0/1); real64 k_arrh(3,nreac), elecenergy(nreac),
!        molmass(nspecies), spec_charge(nspecies); then per state real64
!        rate(nreac) (getarrhrate or getcustomrate at Te and efield_mag),
!        prod(nspecies) (getspecproduction with the field magnitude, as the
!        plasma solver's callers pass it), mob(nspecies) and dco(nspecies)
!        (for the charged species and the solved neutrals, 0 for the
!        others), inelterm (getelectroninelasticterm, with the field),
!        collfreq (getelectroncollisionfrequency).
program ref_chem4
    use chem_module
    implicit none
    character(len=512) :: infile, outfile
    integer :: nst, s, r, j, isr(nreac)
    real*8 :: Tg, Pg, Te, ef, sd(nspecies), rate(nreac), prod(nspecies), mob(nspecies), dco(nspecies)
    real*8 :: inel, cf, p

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
            if (spec_charge(j) .ne. 0.d0) mob(j) = getspecmobility(j, Te, Tg, Pg)
            if (spec_charge(j) .ne. 0.d0 .or. (j .ge. neutralspecmin .and. j .le. neutralspecmax)) &
                dco(j) = getspecdcoeff(j, Te, Tg, Pg)
        enddo
        call getelectroninelasticterm(Te, Tg, sd, inel, ef)
        cf = getelectroncollisionfrequency(Te, Tg, Pg)
        write(22) rate, prod, mob, dco, inel, cf
    enddo
    close(21)
    close(22)
end program ref_chem4
