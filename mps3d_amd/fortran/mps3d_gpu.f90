! ISO_C_BINDING interface to the mps3d_amd HIP library (include/mps3d.h).
!
! The reference's equation modules call perform_vcycle
! (mgridlib/mgridsteps.f90:384-385) in a loop of at most nvcycles cycles
! (plasma_pde/poissoneq.f90:427-450, generic_pde/pdedefs.f90:378-402).
! gpu_perform_vcycle keeps that argument list -- with the six bcvals faces
! passed as arrays, since the boundarycondition type of
! mgridlib/globalvars.f90:57-66 lives in the reference tree -- and
! gpu_solve replaces the whole loop with one device-resident call.
! Solution and coefficient arrays are the reference's ghosted real*8
! arrays (0:lx+1, 0:ly+1, 0:lz+1); they are passed by address, unchanged.
module mps3d_gpu

    use iso_c_binding
    implicit none

    integer(c_int), parameter :: MPS_OK = 0, MPS_NOT_CONVERGED = 1, MPS_ERR_NAN = -1
    integer(c_int), parameter :: MPS_ERR_ARG = -2, MPS_ERR_HIP = -3, MPS_ERR_STATE = -4, MPS_ERR_NOLIB = -5
    integer(c_int), parameter :: MPS_HIER_REF = 0, MPS_HIER_ALIGNED = 1
    integer(c_int), parameter :: MPS_PREC_FP64 = 0, MPS_PREC_FP32_DC = 1
    integer(c_int), parameter :: MPS_FLUX_UPWIND = 0, MPS_FLUX_SG = 1
    integer(c_int), parameter :: MPS_TOL_RMS = 0, MPS_TOL_REL_L2 = 1
    integer(c_int), parameter :: MPS_ARITH_EXACT = 0, MPS_ARITH_FAST = 1
    integer(c_int), parameter :: MPS_START_GUESS = 0, MPS_START_FMG = 1

    ! chemistry tables (mps_chem): the arrays of a chem_module, by address
    type, bind(C) :: mps_chem
        integer(c_int) :: nspecies, nreac, especnum, bgspecnum, ionspecmin, ionspecmax
        type(c_ptr) :: reactants, products, k_arrh, elecenergy, mobility, dcoeff, molmass, charge
        type(c_ptr) :: mob_kind = c_null_ptr, mob_terms = c_null_ptr, mob_coef = c_null_ptr
        type(c_ptr) :: dcoeff_kind = c_null_ptr
        type(c_ptr) :: rate_kind = c_null_ptr, rate_coef = c_null_ptr
    end type mps_chem

    type, bind(C) :: mps_grid
        integer(c_int) :: nx, ny, nz
        real(c_double) :: lenx, leny, lenz
        integer(c_int) :: px, py, pz
    end type mps_grid

    type, bind(C) :: mps_opts
        integer(c_int) :: hierarchy, precision, flux, nu1, nu2, max_cycles
        real(c_double) :: tol
        integer(c_int) :: tol_kind, arith, start, dc_inner
        real(c_double) :: omega
    end type mps_opts

    interface
        subroutine mps_default_opts(opts) bind(C, name='mps_default_opts')
            import :: mps_opts
            type(mps_opts), intent(out) :: opts
        end subroutine mps_default_opts

        integer(c_int) function mps_create(ctx, grid, opts) bind(C, name='mps_create')
            import :: c_int, c_ptr, mps_grid, mps_opts
            type(c_ptr), intent(out) :: ctx
            type(mps_grid), intent(in) :: grid
            type(mps_opts), intent(in) :: opts
        end function mps_create

        integer(c_int) function mps_destroy(ctx) bind(C, name='mps_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function mps_destroy

        integer(c_int) function mps_get_grid(ctx, grid, nspecies) bind(C, name='mps_get_grid')
            import :: c_int, c_ptr, mps_grid
            type(c_ptr), value :: ctx
            type(mps_grid), intent(out) :: grid
            integer(c_int), intent(out) :: nspecies
        end function mps_get_grid

        integer(c_int) function mps_set_bc_str(ctx, codes, faces) bind(C, name='mps_set_bc_str')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), dimension(*), intent(in) :: codes
            type(c_ptr), dimension(6), intent(in) :: faces
        end function mps_set_bc_str

        integer(c_int) function mps_set_coeffs(ctx, dcoeff, velx, vely, velz, reac, source, timederiv, dt) &
                bind(C, name='mps_set_coeffs')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, dcoeff, velx, vely, velz, reac, source
            integer(c_int), value :: timederiv
            real(c_double), value :: dt
        end function mps_set_coeffs

        integer(c_int) function mps_vcycle(ctx, soln, mgridsterm, resnorm) bind(C, name='mps_vcycle')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, soln, mgridsterm
            real(c_double), intent(out) :: resnorm
        end function mps_vcycle

        integer(c_int) function mps_solve(ctx, soln, ncycles, resnorm, history) bind(C, name='mps_solve')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, soln, history
            integer(c_int), intent(out) :: ncycles
            real(c_double), intent(out) :: resnorm
        end function mps_solve

        integer(c_int) function mps_solve_rhs(ctx, b, soln, ncycles, resnorm, history) &
                bind(C, name='mps_solve_rhs')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, b, soln, history
            integer(c_int), intent(out) :: ncycles
            real(c_double), intent(out) :: resnorm
        end function mps_solve_rhs

        integer(c_int) function mps_efield(ctx, phi, efield, scale, limiter) bind(C, name='mps_efield')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, phi, efield
            real(c_double), value :: scale
            integer(c_int), value :: limiter
        end function mps_efield

        integer(c_int) function mps_set_chemistry(ctx, chem) bind(C, name='mps_set_chemistry')
            import :: c_int, c_ptr, mps_chem
            type(c_ptr), value :: ctx
            type(mps_chem), intent(in) :: chem
        end function mps_set_chemistry

        integer(c_int) function mps_species_transport(ctx, specnum, dt, timederiv, efield, electemp, numden, &
                Tg, Pg, scaling, soln, vel, dcoeff, reac, source, prod) bind(C, name='mps_species_transport')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, efield, electemp, numden, scaling, soln, vel, dcoeff, reac, source, prod
            integer(c_int), value :: specnum, timederiv
            real(c_double), value :: dt, Tg, Pg
        end function mps_species_transport

        integer(c_int) function mps_eenergy_transport(ctx, dt, timederiv, lscale, efield, electemp, numden, &
                Tg, Pg, scaling, soln, vel, dcoeff, reac, source) bind(C, name='mps_eenergy_transport')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, efield, electemp, numden, scaling, soln, vel, dcoeff, reac, source
            integer(c_int), value :: timederiv
            real(c_double), value :: dt, lscale, Tg, Pg
        end function mps_eenergy_transport

        integer(c_int) function mps_poisson_transport(ctx, dt, timederiv, elecden, iondens, scaling, soln, &
                vel, dcoeff, reac, source) bind(C, name='mps_poisson_transport')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, elecden, iondens, scaling, soln, vel, dcoeff, reac, source
            integer(c_int), value :: timederiv
            real(c_double), value :: dt
        end function mps_poisson_transport

        integer(c_int) function mps_electron_temp(ctx, eenergysoln, elecden, eenergyscale, etempmin, electemp) &
                bind(C, name='mps_electron_temp')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, eenergysoln, elecden, electemp
            real(c_double), value :: eenergyscale, etempmin
        end function mps_electron_temp

        integer(c_int) function mps_ion_drift_faces(ctx, specnum, efield, numden, ionsoln, velscale, faces) &
                bind(C, name='mps_ion_drift_faces')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, efield, numden, ionsoln
            integer(c_int), value :: specnum
            real(c_double), value :: velscale
            type(c_ptr), intent(in) :: faces(6)
        end function mps_ion_drift_faces

        integer(c_int) function mps_fill_ghosts(ctx, soln) bind(C, name='mps_fill_ghosts')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, soln
        end function mps_fill_ghosts

        type(c_ptr) function mps_last_error_c() bind(C, name='mps_last_error')
            import :: c_ptr
        end function mps_last_error_c
    end interface

contains

    ! MPS_ERR_ARG unless (lx, ly, lz) is the context's grid and, with
    ! nsp >= 0, nsp the number of species of its chemistry: the wrappers'
    ! array arguments are sized from these, and the library reads and
    ! writes whole arrays of the context's grid
    integer function gpu_check_grid(ctx, lx, ly, lz, nsp) result(status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz, nsp
        type(mps_grid) :: g
        integer(c_int) :: ns
        status = mps_get_grid(ctx, g, ns)
        if (status /= MPS_OK) return
        if (g%nx /= lx .or. g%ny /= ly .or. g%nz /= lz) status = MPS_ERR_ARG
        if (nsp >= 0 .and. nsp /= ns) status = MPS_ERR_ARG
    end function gpu_check_grid

    ! context for the grid and process layout of the reference's globals
    ! (g_Nx.., g_xlen.., g_px.. of mgridlib/globalvars.f90:33-52)
    integer function gpu_open(ctx, nx, ny, nz, xlen, ylen, zlen, px, py, pz, opts) result(status)
        type(c_ptr), intent(out) :: ctx
        integer, intent(in) :: nx, ny, nz, px, py, pz
        real*8, intent(in) :: xlen, ylen, zlen
        type(mps_opts), intent(in) :: opts
        type(mps_grid) :: g
        g%nx = nx
        g%ny = ny
        g%nz = nz
        g%lenx = xlen
        g%leny = ylen
        g%lenz = zlen
        g%px = px
        g%py = py
        g%pz = pz
        status = mps_create(ctx, g, opts)
    end function gpu_open

    ! boundary codes and faces, coefficients (as perform_vcycle receives them)
    integer function gpu_set_problem(ctx, lx, ly, lz, timederivflag, dt, vel, dcoeff, reac, source, &
            bc_codes, left, right, bottom, top, back, front) result(status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz
        logical, intent(in) :: timederivflag
        real*8, intent(in) :: dt
        real*8, intent(in), target :: vel(0:lx+1, 0:ly+1, 0:lz+1, 3)
        real*8, intent(in), target :: dcoeff(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: reac(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: source(0:lx+1, 0:ly+1, 0:lz+1)
        character(len=4), intent(in) :: bc_codes(6)
        real*8, intent(in), target :: left(ly, lz), right(ly, lz), bottom(lx, lz), top(lx, lz)
        real*8, intent(in), target :: back(lx, ly), front(lx, ly)
        character(kind=c_char, len=24) :: codes
        type(c_ptr) :: faces(6)
        integer :: f
        integer(c_int) :: td
        status = gpu_check_grid(ctx, lx, ly, lz, -1)
        if (status /= MPS_OK) return
        do f = 1, 6
            codes(4*f-3:4*f) = bc_codes(f)
        enddo
        faces(1) = c_loc(left)
        faces(2) = c_loc(right)
        faces(3) = c_loc(bottom)
        faces(4) = c_loc(top)
        faces(5) = c_loc(back)
        faces(6) = c_loc(front)
        status = mps_set_bc_str(ctx, codes, faces)
        if (status /= MPS_OK) return
        td = 0
        if (timederivflag) td = 1
        status = mps_set_coeffs(ctx, c_loc(dcoeff), c_loc(vel(0, 0, 0, 1)), c_loc(vel(0, 0, 0, 2)), &
            c_loc(vel(0, 0, 0, 3)), c_loc(reac), c_loc(source), td, dt)
    end function gpu_set_problem

    ! perform_vcycle (mgridlib/mgridsteps.f90:384-579) on the finest level,
    ! argument for argument; the coefficients are set up on every call, as
    ! the reference restricts them on every call
    subroutine gpu_perform_vcycle(ctx, solnvar, timederivflag, dt, lx, ly, lz, vel, dcoeff, reac, source, &
            bc_codes, left, right, bottom, top, back, front, mgridsterm, resnorm, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz
        real*8, intent(inout), target :: solnvar(0:lx+1, 0:ly+1, 0:lz+1)
        logical, intent(in) :: timederivflag
        real*8, intent(in) :: dt
        real*8, intent(in), target :: vel(0:lx+1, 0:ly+1, 0:lz+1, 3)
        real*8, intent(in), target :: dcoeff(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: reac(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: source(0:lx+1, 0:ly+1, 0:lz+1)
        character(len=4), intent(in) :: bc_codes(6)
        real*8, intent(in), target :: left(ly, lz), right(ly, lz), bottom(lx, lz), top(lx, lz)
        real*8, intent(in), target :: back(lx, ly), front(lx, ly)
        real*8, intent(in), target :: mgridsterm(lx, ly, lz)
        real*8, intent(out) :: resnorm
        integer, intent(out) :: status
        status = gpu_set_problem(ctx, lx, ly, lz, timederivflag, dt, vel, dcoeff, reac, source, &
            bc_codes, left, right, bottom, top, back, front)
        if (status /= MPS_OK) return
        status = mps_vcycle(ctx, c_loc(solnvar), c_loc(mgridsterm), resnorm)
    end subroutine gpu_perform_vcycle

    ! the equation modules' loop of V-cycles in one call: set the problem
    ! once with gpu_set_problem, then solve to the context's tolerance
    subroutine gpu_solve(ctx, solnvar, lx, ly, lz, ncycles, resnorm, history, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz
        real*8, intent(inout), target :: solnvar(0:lx+1, 0:ly+1, 0:lz+1)
        integer, intent(out) :: ncycles
        real*8, intent(out) :: resnorm
        real*8, intent(out), target :: history(*)
        integer, intent(out) :: status
        integer(c_int) :: n
        status = gpu_check_grid(ctx, lx, ly, lz, -1)
        if (status /= MPS_OK) return
        status = mps_solve(ctx, c_loc(solnvar), n, resnorm, c_loc(history))
        ncycles = n
    end subroutine gpu_solve

    ! the same solve for a right-hand side b (lx,ly,lz) given directly, with
    ! the coefficients of the last gpu_set_problem: for a Poisson solve per
    ! time step whose space charge changes and whose permittivity does not
    subroutine gpu_solve_rhs(ctx, b, solnvar, lx, ly, lz, ncycles, resnorm, history, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz
        real*8, intent(in), target :: b(lx, ly, lz)
        real*8, intent(inout), target :: solnvar(0:lx+1, 0:ly+1, 0:lz+1)
        integer, intent(out) :: ncycles
        real*8, intent(out) :: resnorm
        real*8, intent(out), target :: history(*)
        integer, intent(out) :: status
        integer(c_int) :: n
        status = gpu_check_grid(ctx, lx, ly, lz, -1)
        if (status /= MPS_OK) return
        status = mps_solve_rhs(ctx, c_loc(b), c_loc(solnvar), n, resnorm, c_loc(history))
        ncycles = n
    end subroutine gpu_solve_rhs

    ! update_efield (plasma_pde/poissoneq.f90:565-680) on the device:
    ! poissonsoln's ghosts from the context's boundary conditions, then
    ! efield = -grad(poissonsoln) * potscale/lscale with zero-gradient ghosts
    subroutine gpu_update_efield(ctx, poissonsoln, efield, lx, ly, lz, potscale, lscale, limiter, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz
        real*8, intent(inout), target :: poissonsoln(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(inout), target :: efield(0:lx+1, 0:ly+1, 0:lz+1, 3)
        real*8, intent(in) :: potscale, lscale
        logical, intent(in) :: limiter
        integer, intent(out) :: status
        integer(c_int) :: lim
        status = gpu_check_grid(ctx, lx, ly, lz, -1)
        if (status /= MPS_OK) return
        lim = 0
        if (limiter) lim = 1
        status = mps_efield(ctx, c_loc(poissonsoln), c_loc(efield), potscale / lscale, lim)
    end subroutine gpu_update_efield

    ! initializechemistry: the chem_module's tables (reactants/products
    ! (nreac, nspecies), k_arrh(3, nreac), per-species mobility, dcoeff,
    ! molmass, charge) copied to the device; with the four optional
    ! arguments, field-dependent mobility and diffusion models (mps_chem in
    ! include/mps3d.h)
    subroutine gpu_set_chemistry(ctx, nspecies, nreac, especnum, bgspecnum, ionspecmin, ionspecmax, &
                                 reactants, products, k_arrh, elecenergy, mobility, dcoeff, molmass, charge, status, &
                                 mob_kind, mob_terms, mob_coef, dcoeff_kind)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: nspecies, nreac, especnum, bgspecnum, ionspecmin, ionspecmax
        integer(c_int), intent(in), target :: reactants(nreac, nspecies), products(nreac, nspecies)
        real*8, intent(in), target :: k_arrh(3, nreac), elecenergy(nreac)
        real*8, intent(in), target :: mobility(nspecies), dcoeff(nspecies), molmass(nspecies), charge(nspecies)
        integer, intent(out) :: status
        integer(c_int), intent(in), target, optional :: mob_kind(nspecies), mob_terms(nspecies, nspecies)
        integer(c_int), intent(in), target, optional :: dcoeff_kind(nspecies)
        real*8, intent(in), target, optional :: mob_coef(3, nspecies, nspecies)
        type(mps_chem) :: ch
        ch%nspecies = nspecies
        ch%nreac = nreac
        ch%especnum = especnum
        ch%bgspecnum = bgspecnum
        ch%ionspecmin = ionspecmin
        ch%ionspecmax = ionspecmax
        ch%reactants = c_loc(reactants)
        ch%products = c_loc(products)
        ch%k_arrh = c_loc(k_arrh)
        ch%elecenergy = c_loc(elecenergy)
        ch%mobility = c_loc(mobility)
        ch%dcoeff = c_loc(dcoeff)
        ch%molmass = c_loc(molmass)
        ch%charge = c_loc(charge)
        if (present(mob_kind) .and. present(mob_terms) .and. present(mob_coef) .and. present(dcoeff_kind)) then
            ch%mob_kind = c_loc(mob_kind)
            ch%mob_terms = c_loc(mob_terms)
            ch%mob_coef = c_loc(mob_coef)
            ch%dcoeff_kind = c_loc(dcoeff_kind)
        endif
        status = mps_set_chemistry(ctx, ch)
    end subroutine gpu_set_chemistry

    ! eden_update_transport / ions_update_transport with the module's
    ! vel, dcoeff, reac, source arrays and scalings passed explicitly;
    ! prod = elecprod for the electrons (pass any array for ions: its
    ! interior gets the unscaled production)
    subroutine gpu_species_transport(ctx, specnum, dt, elecfield, electemp, numden, Tg, Pg, scaling, soln, &
                                     vel, dcoeff, reac, source, prod, lx, ly, lz, nsp, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: specnum, lx, ly, lz, nsp
        real*8, intent(in) :: dt, Tg, Pg
        real*8, intent(in), target :: elecfield(0:lx+1, 0:ly+1, 0:lz+1, 3), electemp(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: numden(0:lx+1, 0:ly+1, 0:lz+1, nsp), soln(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: scaling(4)
        real*8, intent(inout), target :: vel(0:lx+1, 0:ly+1, 0:lz+1, 3), dcoeff(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(inout), target :: reac(0:lx+1, 0:ly+1, 0:lz+1), source(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(inout), target :: prod(0:lx+1, 0:ly+1, 0:lz+1)
        integer, intent(out) :: status
        status = gpu_check_grid(ctx, lx, ly, lz, nsp)
        if (status /= MPS_OK) return
        status = mps_species_transport(ctx, specnum, dt, 1, c_loc(elecfield), c_loc(electemp), c_loc(numden), &
            Tg, Pg, c_loc(scaling), c_loc(soln), c_loc(vel), c_loc(dcoeff), c_loc(reac), c_loc(source), c_loc(prod))
    end subroutine gpu_species_transport

    ! eenergy_update_transport (lscale: the length scale of the Joule term's dx)
    subroutine gpu_eenergy_transport(ctx, dt, lscale, elecfield, electemp, numden, Tg, Pg, scaling, soln, &
                                     vel, dcoeff, reac, source, lx, ly, lz, nsp, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz, nsp
        real*8, intent(in) :: dt, lscale, Tg, Pg
        real*8, intent(in), target :: elecfield(0:lx+1, 0:ly+1, 0:lz+1, 3), electemp(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: numden(0:lx+1, 0:ly+1, 0:lz+1, nsp), soln(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: scaling(4)
        real*8, intent(inout), target :: vel(0:lx+1, 0:ly+1, 0:lz+1, 3), dcoeff(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(inout), target :: reac(0:lx+1, 0:ly+1, 0:lz+1), source(0:lx+1, 0:ly+1, 0:lz+1)
        integer, intent(out) :: status
        status = gpu_check_grid(ctx, lx, ly, lz, nsp)
        if (status /= MPS_OK) return
        status = mps_eenergy_transport(ctx, dt, 1, lscale, c_loc(elecfield), c_loc(electemp), c_loc(numden), &
            Tg, Pg, c_loc(scaling), c_loc(soln), c_loc(vel), c_loc(dcoeff), c_loc(reac), c_loc(source))
    end subroutine gpu_eenergy_transport

    ! poisson_update_transport (iondens: the solved ions one after the other)
    subroutine gpu_poisson_transport(ctx, dt, elecden, iondens, scaling, soln, vel, dcoeff, reac, source, &
                                     lx, ly, lz, nions, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz, nions
        real*8, intent(in) :: dt
        real*8, intent(in), target :: elecden(0:lx+1, 0:ly+1, 0:lz+1), iondens(0:lx+1, 0:ly+1, 0:lz+1, nions)
        real*8, intent(in), target :: scaling(4), soln(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(inout), target :: vel(0:lx+1, 0:ly+1, 0:lz+1, 3), dcoeff(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(inout), target :: reac(0:lx+1, 0:ly+1, 0:lz+1), source(0:lx+1, 0:ly+1, 0:lz+1)
        integer, intent(out) :: status
        status = gpu_check_grid(ctx, lx, ly, lz, -1)
        if (status /= MPS_OK) return
        status = mps_poisson_transport(ctx, dt, 1, c_loc(elecden), c_loc(iondens), c_loc(scaling), c_loc(soln), &
            c_loc(vel), c_loc(dcoeff), c_loc(reac), c_loc(source))
    end subroutine gpu_poisson_transport

    ! update_electrontemp
    subroutine gpu_update_electrontemp(ctx, electemp, elecden, eenergysoln, eenergyscale, etempmin, &
                                       lx, ly, lz, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: lx, ly, lz
        real*8, intent(inout), target :: electemp(0:lx+1, 0:ly+1, 0:lz+1), eenergysoln(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in), target :: elecden(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in) :: eenergyscale, etempmin
        integer, intent(out) :: status
        status = gpu_check_grid(ctx, lx, ly, lz, -1)
        if (status /= MPS_OK) return
        status = mps_electron_temp(ctx, c_loc(eenergysoln), c_loc(elecden), eenergyscale, etempmin, c_loc(electemp))
    end subroutine gpu_update_electrontemp

    ! the FLUX faces of ions_update_bcs for ion specnum: the faces with
    ! flux(f) /= 0 (left, right, bottom, top, back, front) get the drift outflow
    subroutine gpu_ions_flux_faces(ctx, specnum, efield, numden, ionsoln, velscale, flux, leftface, rightface, &
                                   bottomface, topface, backface, frontface, lx, ly, lz, nspec, status)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: specnum, lx, ly, lz, nspec, flux(6)
        real*8, intent(in), target :: efield(0:lx+1, 0:ly+1, 0:lz+1, 3)
        real*8, intent(in), target :: numden(0:lx+1, 0:ly+1, 0:lz+1, nspec), ionsoln(0:lx+1, 0:ly+1, 0:lz+1)
        real*8, intent(in) :: velscale
        real*8, intent(inout), target :: leftface(ly, lz), rightface(ly, lz), bottomface(lx, lz), topface(lx, lz)
        real*8, intent(inout), target :: backface(lx, ly), frontface(lx, ly)
        integer, intent(out) :: status
        type(c_ptr) :: f(6)
        status = gpu_check_grid(ctx, lx, ly, lz, nspec)
        if (status /= MPS_OK) return
        f = c_null_ptr
        if (flux(1) /= 0) f(1) = c_loc(leftface)
        if (flux(2) /= 0) f(2) = c_loc(rightface)
        if (flux(3) /= 0) f(3) = c_loc(bottomface)
        if (flux(4) /= 0) f(4) = c_loc(topface)
        if (flux(5) /= 0) f(5) = c_loc(backface)
        if (flux(6) /= 0) f(6) = c_loc(frontface)
        status = mps_ion_drift_faces(ctx, int(specnum, c_int), c_loc(efield), c_loc(numden), c_loc(ionsoln), &
                                     velscale, f)
    end subroutine gpu_ions_flux_faces

    ! text of the library's last error
    function gpu_last_error() result(msg)
        character(len=256) :: msg
        character(kind=c_char), pointer :: p(:)
        integer :: k
        call c_f_pointer(mps_last_error_c(), p, [256])
        msg = ' '
        do k = 1, 256
            if (p(k) == c_null_char) exit
            msg(k:k) = p(k)
        enddo
    end function gpu_last_error

end module mps3d_gpu
