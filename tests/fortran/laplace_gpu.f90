! Fortran host driving the HIP library through ISO_C_BINDING
! (mps3d_amd/fortran/mps3d_gpu.f90), on the Laplace deck of
! tests/Laplace_test (INPUTS:13-27, SOLVER:4-7): 128^3 cells on a unit cube,
! 2x2x2 process grid, phi = 0 / 1 on the x faces (DIRC), zero gradient
! elsewhere, up to 20 V-cycles. The exact discrete solution is phi = x at
! cell centres (tests/Laplace_test/README:1-4).
!
! Part 1 loops gpu_perform_vcycle the way generic_pde/pdedefs.f90:378-402
! loops perform_vcycle; part 2 makes the same solve with one gpu_solve;
! part 3 with gpu_solve_rhs and the right-hand side passed directly;
! part 4 takes the field of that potential with gpu_update_efield, as
! plasma_pde/poissoneq.f90:549-680 does after each solve: for phi = x and
! potscale / lscale = 2, E = (-2, 0, 0) in every cell and ghost face;
! part 5 sets the ion_elec chemistry (plasma_pde/chemistry/ion_elec_chem.f90)
! and assembles the electron transport in that field (vel = mu_e E) and
! the Poisson space charge e/eps0 (n_i - n_e) on the device.
! Exit status: 0 pass, 1 fail, 2 no usable GPU.
program laplace_gpu
    use mps3d_gpu
    implicit none
    integer, parameter :: n = 128, nvcycles = 20
    real*8, allocatable, target :: u(:,:,:), ef(:,:,:,:), vel(:,:,:,:), tv(:,:,:,:), nd(:,:,:,:)
    real*8, allocatable, target :: td(:,:,:), tr(:,:,:), ts(:,:,:), tp(:,:,:), te(:,:,:)
    integer(c_int), target :: reactants(1, 3), products(1, 3)
    real*8, allocatable, target :: dc(:,:,:), reac(:,:,:), src(:,:,:), sterm(:,:,:)
    real*8, target :: k_arrh(3, 1), elecenergy(1), mob(3), dco(3), mass(3), chg(3), scal(4)
    real*8, allocatable, target :: left(:,:), right(:,:), bottom(:,:), top(:,:), back(:,:), front(:,:)
    real*8, target :: hist(nvcycles)
    character(len=4) :: codes(6)
    type(c_ptr) :: ctx
    type(mps_opts) :: opts
    real*8 :: resnorm, err, x, tol
    integer :: i, it, st, ncyc

    allocate(u(0:n+1, 0:n+1, 0:n+1), ef(0:n+1, 0:n+1, 0:n+1, 3), vel(0:n+1, 0:n+1, 0:n+1, 3), dc(0:n+1, 0:n+1, 0:n+1))
    allocate(reac(0:n+1, 0:n+1, 0:n+1), src(0:n+1, 0:n+1, 0:n+1), sterm(n, n, n))
    allocate(left(n, n), right(n, n), bottom(n, n), top(n, n), back(n, n), front(n, n))
    u = 0.d0
    vel = 0.d0
    dc = 1.d0
    reac = 0.d0
    src = 0.d0
    sterm = 0.d0
    left = 0.d0
    right = 1.d0
    bottom = 0.d0
    top = 0.d0
    back = 0.d0
    front = 0.d0
    codes = (/ 'DIRC', 'DIRC', 'ZGRD', 'ZGRD', 'ZGRD', 'ZGRD' /)
    tol = 1.d-11 * 0.001   ! ERRTOLR*0.001 as generic_pde/pdedefs.f90:368

    call mps_default_opts(opts)
    opts%max_cycles = nvcycles
    opts%tol = tol
    st = gpu_open(ctx, n, n, n, 1.d0, 1.d0, 1.d0, 2, 2, 2, opts)
    if (st == MPS_ERR_NOLIB) then
        print *, 'laplace_gpu: no usable GPU'
        call exit(2)
    endif
    if (st /= MPS_OK) then
        print *, 'laplace_gpu: create failed ', st, trim(gpu_last_error())
        call exit(1)
    endif

    ! part 1: the reference's loop, one perform_vcycle per iteration
    do it = 1, nvcycles
        sterm = 0.d0
        call gpu_perform_vcycle(ctx, u, .false., 0.1d0, n, n, n, vel, dc, reac, src, codes, &
            left, right, bottom, top, back, front, sterm, resnorm, st)
        if (st /= MPS_OK) then
            print *, 'laplace_gpu: perform_vcycle failed ', st, trim(gpu_last_error())
            call exit(1)
        endif
        print '(a,i3,es14.6)', ' it:', it, resnorm
        if (resnorm <= tol) exit
    enddo
    err = 0.d0
    do i = 1, n
        x = (i - 0.5d0) / n
        err = max(err, maxval(abs(u(i, 1:n, 1:n) - x)))
    enddo
    print '(a,es12.4)', ' loop of gpu_perform_vcycle: max |phi - x| =', err
    if (err > 1.d-10) call exit(1)

    ! part 2: the whole loop in one call
    u = 0.d0
    call gpu_solve(ctx, u, n, n, n, ncyc, resnorm, hist, st)
    if (st /= MPS_OK .and. st /= MPS_NOT_CONVERGED) then
        print *, 'laplace_gpu: solve failed ', st, trim(gpu_last_error())
        call exit(1)
    endif
    err = 0.d0
    do i = 1, n
        x = (i - 0.5d0) / n
        err = max(err, maxval(abs(u(i, 1:n, 1:n) - x)))
    enddo
    print '(a,i3,a,es12.4,a,es12.4)', ' gpu_solve: cycles', ncyc, ' resnorm', resnorm, ' max |phi - x| =', err
    if (err > 1.d-10) call exit(1)

    ! part 3: the same system with its right-hand side passed directly (what
    ! find_bvec returns for the deck: the DIRC term 2 D g / h^2 of the x = 1
    ! face, D = g = 1, h = 1/n), the coefficients as set above
    sterm = 0.d0
    sterm(n, :, :) = 2.d0 * dble(n) * dble(n)
    u = 0.d0
    call gpu_solve_rhs(ctx, sterm, u, n, n, n, ncyc, resnorm, hist, st)
    if (st /= MPS_OK .and. st /= MPS_NOT_CONVERGED) then
        print *, 'laplace_gpu: solve_rhs failed ', st, trim(gpu_last_error())
        call exit(1)
    endif
    err = 0.d0
    do i = 1, n
        x = (i - 0.5d0) / n
        err = max(err, maxval(abs(u(i, 1:n, 1:n) - x)))
    enddo
    print '(a,i3,a,es12.4,a,es12.4)', ' gpu_solve_rhs: cycles', ncyc, ' resnorm', resnorm, ' max |phi - x| =', err
    if (err > 1.d-10) call exit(1)

    ! part 4: E = -(potscale / lscale) grad phi, ghost faces by zero gradient
    ef = 0.d0
    call gpu_update_efield(ctx, u, ef, n, n, n, 4.d0, 2.d0, .false., st)
    if (st /= MPS_OK) then
        print *, 'laplace_gpu: update_efield failed ', st, trim(gpu_last_error())
        call exit(1)
    endif
    err = max(maxval(abs(ef(0:n+1, 1:n, 1:n, 1) + 2.d0)), maxval(abs(ef(1:n, 0:n+1, 1:n, 2))), &
              maxval(abs(ef(1:n, 1:n, 0:n+1, 3))))
    print '(a,es12.4)', ' gpu_update_efield: max |E - (-2, 0, 0)| =', err
    if (err > 1.d-6) call exit(1)

    ! part 5: ion_elec chemistry, electron transport and Poisson space charge
    reactants = 0
    products = 0
    reactants(1, 1) = 1
    reactants(1, 2) = 1
    products(1, 3) = 1
    products(1, 2) = 2
    k_arrh(:, 1) = (/ 2.5d-6 * 0.01**3, 0.d0, 278508.d0 /)
    elecenergy = -15.578
    mob = (/ 0.d0, -20.d0, 0.2d0 /)
    dco = (/ 0.d0, 100.d0, 0.01d0 /)
    mass = (/ 4.d0 * 1.672621d-27, 9.109382d-31, 4.d0 * 1.672621d-27 /)
    chg = (/ 0.d0, -1.d0, 1.d0 /)
    call gpu_set_chemistry(ctx, 3, 1, 2, 1, 3, 3, reactants, products, k_arrh, elecenergy, mob, dco, mass, chg, st)
    if (st /= MPS_OK) then
        print *, 'laplace_gpu: set_chemistry failed ', st, trim(gpu_last_error())
        call exit(1)
    endif
    allocate(tv(0:n+1, 0:n+1, 0:n+1, 3), nd(0:n+1, 0:n+1, 0:n+1, 3), te(0:n+1, 0:n+1, 0:n+1))
    allocate(td(0:n+1, 0:n+1, 0:n+1), tr(0:n+1, 0:n+1, 0:n+1), ts(0:n+1, 0:n+1, 0:n+1), tp(0:n+1, 0:n+1, 0:n+1))
    nd(:, :, :, 1) = 1.d22
    nd(:, :, :, 2) = 1.d15
    nd(:, :, :, 3) = 3.d15
    te = 2.d4
    tp = 0.d0
    scal = (/ 0.5d0, 0.25d0, 1.d0, 2.d0 /)
    call gpu_species_transport(ctx, 2, 1.d0, ef, te, nd, 300.d0, 101325.d0, scal, u, tv, td, tr, ts, tp, &
                               n, n, n, 3, st)
    if (st /= MPS_OK) then
        print *, 'laplace_gpu: species_transport failed ', st, trim(gpu_last_error())
        call exit(1)
    endif
    ! mu_e E_x vel_scaling = -20 * (-2) * 0.5 up to the solve's error in E
    err = max(maxval(abs(tv(1:n, 1:n, 1:n, 1) - 20.d0)), maxval(abs(td(1:n, 1:n, 1:n) - 25.d0)))
    print '(a,es12.4)', ' gpu_species_transport: max |vel_x - 20|, |dcoeff - 25| =', err
    if (err > 1.d-5 .or. maxval(tp(1:n, 1:n, 1:n)) <= 0.d0) call exit(1)
    call gpu_poisson_transport(ctx, 1.d0, nd(:, :, :, 2), nd(:, :, :, 3:3), scal, u, tv, td, tr, ts, n, n, n, 1, st)
    if (st /= MPS_OK) then
        print *, 'laplace_gpu: poisson_transport failed ', st, trim(gpu_last_error())
        call exit(1)
    endif
    err = 0.d0
    do i = 1, n
        err = max(err, maxval(abs(ts(i, 1:n, 1:n) - (2.d15 * 1.602176d-19 / 8.854188d-12 * 2.d0 + u(i, 1:n, 1:n))) &
                  / (2.d15 * 1.602176d-19 / 8.854188d-12 * 2.d0)))
    enddo
    print '(a,es12.4)', ' gpu_poisson_transport: max rel. error of the space charge =', err
    if (err > 1.d-14) call exit(1)
    st = mps_destroy(ctx)
    print *, 'laplace_gpu: PASS'
end program laplace_gpu
