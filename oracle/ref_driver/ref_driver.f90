! Golden-vector driver for the mps3d reference multigrid library.
!
! This program is test infrastructure. It is compiled here, against the
! reference's own mgridlib sources (by path, never copied), by
! oracle/Makefile into oracle/_ref/ref_driver, and is run only to write the
! fixtures under tests/golden/ and to time the reference on the host CPU.
!
! It plays the role of generic_pde/pdedefs.f90:280-333 (pde_solve) and
! generic_pde/solve_manager.f90:23-42 (solversetup): it fills the g_*
! globals of mgridlib/globalvars.f90:33-52, calls domaindecompose
! (mgridlib/par_decompose.f90:7-85), slices global input arrays onto each
! rank and then calls one of the reference entry points:
!
!   mode 1  find_AX + find_bvec          (mgridlib/discretepde.f90:672, :382)
!   mode 2  perform_gseidel_smoothing    (mgridlib/discretepde.f90:132)
!   mode 3  loop of perform_vcycle       (mgridlib/mgridsteps.f90:384; the
!           loop of generic_pde/pdedefs.f90:378-402)
!   mode 4  performgmres with find_AX and gseidel_precond
!           (mgridlib/gmres.f90:197; as called by mgridsteps.f90:350-355)
!   mode 5  restriction, restrict_boundary, prolongation1
!           (mgridlib/mgridsteps.f90:11, :55, :170)
!   mode 6  update_boundary_ghostvalues  (mgridlib/discretepde.f90:1121)
!   mode 7  writeoutput                  (mgridlib/cartwrite.f90:455-481):
!           u0 and dcoeff as two cell scalars, written to vtkout.pvtr and
!           vtkout/file_<rank>.vtr in the working directory
!
! Input  (stream, little endian): int32 hdr(16), real64 rhdr(5),
!        character(24) bc codes, the six global BC face arrays, then the
!        global interior arrays u0, dcoeff, velx, vely, velz, reac, source,
!        mgridsterm, bvec (each Nx*Ny*Nz), then for modes 5 and 6 one
!        ghosted array.
! Output (stream): records of character(8) name, int64 n, real64 data(n).
program ref_driver

    use globalvars
    use par_decompose
    use discretepde
    use mgridsteps
    use solvergmres_module
    use cartwrite
    implicit none

    character(len=512) :: infile, outfile
    integer :: hdr(16)
    real*8  :: rhdr(5)
    character(len=24) :: codestr
    character(len=4)  :: bc_codes(NFACES)
    integer :: mode, timederiv_i, nvcycles, niter, maxkspdim, nrestarts
    logical :: timederivflag, success
    real*8  :: dt, tol, resnorm, t0, t1, initres, finalres
    integer :: ierr, f, i, j, k, it, nx, ny, nz, lx, ly, lz, ox, oy, oz

    real*8, allocatable :: gleft(:,:), gright(:,:), gbottom(:,:), gtop(:,:), gback(:,:), gfront(:,:)
    real*8, allocatable :: gu0(:,:,:), gd(:,:,:), gvx(:,:,:), gvy(:,:,:), gvz(:,:,:)
    real*8, allocatable :: greac(:,:,:), gsrc(:,:,:), gsterm(:,:,:), gb(:,:,:), gextra(:,:,:)
    real*8, allocatable :: gout(:,:,:)

    real*8, allocatable :: soln(:,:,:), dcoeff(:,:,:), vel(:,:,:,:), reac(:,:,:), source(:,:,:)
    real*8, allocatable :: sterm(:,:,:), bvec(:,:,:), ax(:,:,:), diag(:,:,:), offdiag(:,:,:), res(:,:,:)
    real*8, allocatable :: x0(:,:,:), xs(:,:,:), coarse(:,:,:), fine(:,:,:), hist(:)
    type(boundarycondition) :: bcvals, bcvals_c
    real*8, allocatable :: vtkvars(:,:,:,:)
    character(len=40) :: vtknames(2)

    call MPI_Init(ierr)
    call get_command_argument(1, infile)
    call get_command_argument(2, outfile)

    f = 21
    open(unit=f, file=trim(infile), access='stream', form='unformatted', status='old')
    read(f) hdr
    read(f) rhdr
    read(f) codestr
    do i = 1, NFACES
        bc_codes(i) = codestr(4*i-3:4*i)
    enddo

    mode        = hdr(2)
    g_Nx        = hdr(3)
    g_Ny        = hdr(4)
    g_Nz        = hdr(5)
    g_px        = hdr(6)
    g_py        = hdr(7)
    g_pz        = hdr(8)
    timederiv_i = hdr(9)
    nvcycles    = hdr(10)
    niter       = hdr(11)
    maxkspdim   = hdr(12)
    nrestarts   = hdr(13)
    timederivflag = (timederiv_i .ne. 0)

    g_xlen = rhdr(1)
    g_ylen = rhdr(2)
    g_zlen = rhdr(3)
    dt     = rhdr(4)
    tol    = rhdr(5)
    g_xorigin = ZERO
    g_yorigin = ZERO
    g_zorigin = ZERO

    nx = g_Nx
    ny = g_Ny
    nz = g_Nz

    allocate(gleft(ny,nz), gright(ny,nz), gbottom(nx,nz), gtop(nx,nz), gback(nx,ny), gfront(nx,ny))
    read(f) gleft
    read(f) gright
    read(f) gbottom
    read(f) gtop
    read(f) gback
    read(f) gfront

    allocate(gu0(nx,ny,nz), gd(nx,ny,nz), gvx(nx,ny,nz), gvy(nx,ny,nz), gvz(nx,ny,nz))
    allocate(greac(nx,ny,nz), gsrc(nx,ny,nz), gsterm(nx,ny,nz), gb(nx,ny,nz))
    read(f) gu0
    read(f) gd
    read(f) gvx
    read(f) gvy
    read(f) gvz
    read(f) greac
    read(f) gsrc
    read(f) gsterm
    read(f) gb
    if (mode .eq. 5) then
        allocate(gextra(0:nx/2+1, 0:ny/2+1, 0:nz/2+1))
        read(f) gextra
    else if (mode .eq. 6) then
        allocate(gextra(0:nx+1, 0:ny+1, 0:nz+1))
        read(f) gextra
    endif
    close(f)

    call domaindecompose()

    lx = g_lx
    ly = g_ly
    lz = g_lz
    ox = (g_pindx-1)*lx
    oy = (g_pindy-1)*ly
    oz = (g_pindz-1)*lz

    allocate(soln(0:lx+1,0:ly+1,0:lz+1), dcoeff(0:lx+1,0:ly+1,0:lz+1))
    allocate(vel(0:lx+1,0:ly+1,0:lz+1,NDIM), reac(0:lx+1,0:ly+1,0:lz+1), source(0:lx+1,0:ly+1,0:lz+1))
    allocate(sterm(lx,ly,lz), bvec(lx,ly,lz), ax(lx,ly,lz), diag(lx,ly,lz), offdiag(lx,ly,lz), res(lx,ly,lz))
    soln = ZERO
    dcoeff = ZERO
    vel = ZERO
    reac = ZERO
    source = ZERO
    soln(1:lx,1:ly,1:lz)     = gu0(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    dcoeff(1:lx,1:ly,1:lz)   = gd(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    vel(1:lx,1:ly,1:lz,XDIR) = gvx(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    vel(1:lx,1:ly,1:lz,YDIR) = gvy(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    vel(1:lx,1:ly,1:lz,ZDIR) = gvz(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    reac(1:lx,1:ly,1:lz)     = greac(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    source(1:lx,1:ly,1:lz)   = gsrc(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    sterm                    = gsterm(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
    bvec                     = gb(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)

    call allocate_bcvalues(bcvals, lx, ly, lz)
    bcvals%left   = gleft(oy+1:oy+ly, oz+1:oz+lz)
    bcvals%right  = gright(oy+1:oy+ly, oz+1:oz+lz)
    bcvals%bottom = gbottom(ox+1:ox+lx, oz+1:oz+lz)
    bcvals%top    = gtop(ox+1:ox+lx, oz+1:oz+lz)
    bcvals%back   = gback(ox+1:ox+lx, oy+1:oy+ly)
    bcvals%front  = gfront(ox+1:ox+lx, oy+1:oy+ly)

    if (g_myproc .eq. g_rootproc) then
        open(unit=22, file=trim(outfile), access='stream', form='unformatted', status='replace')
    endif

    select case (mode)
    case (1)
        call find_AX(soln, timederivflag, dt, lx, ly, lz, vel, dcoeff, reac, bc_codes, bcvals, &
            ax, diag, offdiag, g_lrank, g_rrank, g_brank, g_trank, g_krank, g_frank, &
            g_llenx, g_lleny, g_llenz)
        call put_gathered('AX      ', ax)
        call put_gathered('DIAG    ', diag)
        call put_gathered('OFFDIAG ', offdiag)
        call find_bvec(soln, dt, lx, ly, lz, sterm, vel, dcoeff, source, bc_codes, bcvals, &
            .true., bvec, g_lrank, g_rrank, g_brank, g_trank, g_krank, g_frank, &
            g_llenx, g_lleny, g_llenz)
        call put_gathered('BVEC    ', bvec)
        call find_bvec(soln, dt, lx, ly, lz, sterm, vel, dcoeff, source, bc_codes, bcvals, &
            .false., bvec, g_lrank, g_rrank, g_brank, g_trank, g_krank, g_frank, &
            g_llenx, g_lleny, g_llenz)
        call put_gathered('BVECNOBC', bvec)

    case (2)
        call perform_gseidel_smoothing(soln, timederivflag, dt, lx, ly, lz, .true., sterm, res, &
            vel, dcoeff, reac, source, bc_codes, bcvals, &
            g_lrank, g_rrank, g_brank, g_trank, g_krank, g_frank, &
            g_llenx, g_lleny, g_llenz, niter)
        call put_gathered('U       ', soln(1:lx,1:ly,1:lz))
        call put_gathered('RES     ', res)

    case (3)
        allocate(hist(nvcycles))
        hist = -1.d0
        call MPI_Barrier(g_comm, ierr)
        t0 = MPI_Wtime()
        do it = 1, nvcycles
            sterm = gsterm(ox+1:ox+lx, oy+1:oy+ly, oz+1:oz+lz)
            call perform_vcycle(soln, timederivflag, dt, lx, ly, lz, vel, dcoeff, reac, source, &
                bc_codes, bcvals, sterm, resnorm)
            hist(it) = resnorm
            if (resnorm .le. tol) exit
        enddo
        t1 = MPI_Wtime()
        call put_gathered('U       ', soln(1:lx,1:ly,1:lz))
        if (g_myproc .eq. g_rootproc) then
            call put_record('HIST    ', hist, nvcycles)
            call put_record('TIME    ', (/ t1 - t0 /), 1)
        endif

    case (4)
        ! single-rank known-answer run of the Krylov solver on the whole grid
        allocate(x0(0:lx+1,0:ly+1,0:lz+1), xs(0:lx+1,0:ly+1,0:lz+1))
        x0 = soln
        xs = soln
        success = .false.
        call performgmres(bvec, x0, xs, timederivflag, dt, lx, ly, lz, vel, dcoeff, reac, &
            bc_codes, bcvals, g_lrank, g_rrank, g_brank, g_trank, g_krank, g_frank, &
            g_llenx, g_lleny, g_llenz, maxkspdim, nrestarts, find_AX, gseidel_precond, &
            tol, success, .false., initres, finalres)
        call put_gathered('X       ', xs(1:lx,1:ly,1:lz))

    case (5)
        allocate(coarse(lx/2,ly/2,lz/2), fine(0:lx+1,0:ly+1,0:lz+1))
        call restriction(soln(1:lx,1:ly,1:lz), lx, ly, lz, coarse)
        if (g_myproc .eq. g_rootproc) call put_record('RESTR   ', coarse, size(coarse))
        call allocate_bcvalues(bcvals_c, lx/2, ly/2, lz/2)
        call restrict_boundary(bcvals, lx, ly, lz, bcvals_c)
        if (g_myproc .eq. g_rootproc) then
            call put_record('BCLEFT  ', bcvals_c%left, size(bcvals_c%left))
            call put_record('BCRIGHT ', bcvals_c%right, size(bcvals_c%right))
            call put_record('BCBOTTOM', bcvals_c%bottom, size(bcvals_c%bottom))
            call put_record('BCTOP   ', bcvals_c%top, size(bcvals_c%top))
            call put_record('BCBACK  ', bcvals_c%back, size(bcvals_c%back))
            call put_record('BCFRONT ', bcvals_c%front, size(bcvals_c%front))
        endif
        fine = ZERO
        call prolongation1(gextra, lx/2, ly/2, lz/2, fine)
        if (g_myproc .eq. g_rootproc) call put_record('PROLONG ', fine(1:lx,1:ly,1:lz), lx*ly*lz)

    case (6)
        soln = gextra
        call update_boundary_ghostvalues(soln, bc_codes, bcvals, lx, ly, lz)
        if (g_myproc .eq. g_rootproc) call put_record('GHOSTED ', soln, size(soln))

    case (7)
        allocate(vtkvars(0:lx+1,0:ly+1,0:lz+1,2))
        vtkvars = ZERO
        vtkvars(:,:,:,1) = soln
        vtkvars(:,:,:,2) = dcoeff
        vtknames(1) = 'Potential (V)'
        vtknames(2) = 'Electron_density (#/m3)'
        call writeoutput('vtkout', vtkvars, 2, vtknames, lx, ly, lz)
        if (g_myproc .eq. g_rootproc) call put_record('VTKDONE ', (/ 1.d0 /), 1)

    end select

    if (g_myproc .eq. g_rootproc) close(22)
    call MPI_Finalize(ierr)

contains

    subroutine put_record(name, a, n)
        character(len=8), intent(in) :: name
        integer, intent(in) :: n
        real*8, intent(in) :: a(n)
        integer*8 :: n8
        n8 = n
        write(22) name
        write(22) n8
        write(22) a
    end subroutine put_record

    subroutine put_gathered(name, a)
        ! gathers a local interior array onto the root with the reference's
        ! own gatherterms (mgridlib/par_decompose.f90:298-336)
        character(len=8), intent(in) :: name
        real*8, intent(in) :: a(:,:,:)
        real*8, allocatable :: g(:,:,:)
        allocate(g(lx*g_px, ly*g_py, lz*g_pz))
        call gatherterms(a, lx, ly, lz, g)
        if (g_myproc .eq. g_rootproc) call put_record(name, g, size(g))
        deallocate(g)
    end subroutine put_gathered

end program ref_driver
