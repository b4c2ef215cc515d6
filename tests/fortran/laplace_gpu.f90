! Fortran host driving the HIP library through ISO_C_BINDING
! (mps3d_amd/fortran/mps3d_gpu.f90), on the Laplace deck of
! tests/Laplace_test (INPUTS:13-27, SOLVER:4-7): 128^3 cells on a unit cube,
! 2x2x2 process grid, phi = 0 / 1 on the x faces (DIRC), zero gradient
! elsewhere, up to 20 V-cycles. The exact discrete solution is phi = x at
! cell centres (tests/Laplace_test/README:1-4).
!
! Part 1 loops gpu_perform_vcycle the way generic_pde/pdedefs.f90:378-402
! loops perform_vcycle; part 2 makes the same solve with one gpu_solve;
! part 3 with gpu_solve_rhs and the right-hand side passed directly.
! Exit status: 0 pass, 1 fail, 2 no usable GPU.
program laplace_gpu
    use mps3d_gpu
    implicit none
    integer, parameter :: n = 128, nvcycles = 20
    real*8, allocatable, target :: u(:,:,:), vel(:,:,:,:), dc(:,:,:), reac(:,:,:), src(:,:,:), sterm(:,:,:)
    real*8, allocatable, target :: left(:,:), right(:,:), bottom(:,:), top(:,:), back(:,:), front(:,:)
    real*8, target :: hist(nvcycles)
    character(len=4) :: codes(6)
    type(c_ptr) :: ctx
    type(mps_opts) :: opts
    real*8 :: resnorm, err, x, tol
    integer :: i, it, st, ncyc

    allocate(u(0:n+1, 0:n+1, 0:n+1), vel(0:n+1, 0:n+1, 0:n+1, 3), dc(0:n+1, 0:n+1, 0:n+1))
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
    st = mps_destroy(ctx)
    print *, 'laplace_gpu: PASS'
end program laplace_gpu
