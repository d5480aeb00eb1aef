!> One-GPU example driver for the fft_poisson module.
!>
!> Follows the GPU sequence of the reference example (01_Fortran/examples/
!> main.f90:38-61,109-142,175-190): read the namelist input, build the grid,
!> allocate device arrays, create + initialise the plan, load the continue
!> file when ContinueFilein is set (main.f90:136-139), then time the RHS
!> kernel + one cuda_Poisson_FFT_1D + the copy of P back (main.f90:44-61),
!> write the continue file and the Tecplot field, clean up.
!> The reference's MPI process grid collapses to one process: the input's
!> MPI_procs entries are read and ignored (the solution does not depend on
!> the decomposition).
!>
!>   ppf_example [input.dat]        (default ../run/PARA_INPUT.dat)
!>
!> Input: the namelists of 01_Fortran/run/PARA_INPUT_GPU.dat, read by
!> global_inputpara (global.f90:59-120).  Output: timing and the relative FP64
!> residual on stdout, and two checks: the device RHS against the same
!> formula evaluated on the host, and -- with ContinueFilein -- the field read
!> from dir_cont_filein//'cont_P.bin' against the new solution.  With
!> ContinueFileout, dir_cont_fileout//'cont_P.bin'
!> holds the interior P(1:n1m,1:n2m,1:n3m) as raw FP64 in Fortran order, the
!> layout of cuda_Post_FileOut_Continue_Post_Reassembly_IO (cuda_post.f90:260-292);
!> dir_instantfield//'P_000.PLT' is the Tecplot ASCII field of
!> cuda_Post_Fileout_InstantField_3D (cuda_post.f90:294-331): P interpolated
!> to the grid nodes, one "X Y Z P" line per node in E11.4.
program ppf_example
    use iso_fortran_env, only : int64
    use iso_c_binding
    use fft_poisson
    implicit none

    interface
        !> The example's RHS kernel on the device (include/pascal_poisson_fft.h)
        integer(c_int) function c_example_rhs(f, x1, x2, x3, n1m, n2m, n3m, stream) bind(C, name='ppf_example_rhs')
            import :: c_int, c_ptr
            type(c_ptr), value :: f, x1, x2, x3, stream
            integer(c_int), value :: n1m, n2m, n3m
        end function
    end interface

    real(rp) :: L1, L2, L3, gamma1, gamma2, gamma3
    integer :: n1m, n2m, n3m, np1, np2, np3, uniform1, uniform2, uniform3
    logical :: pbc1, pbc2, pbc3, ContinueFilein, ContinueFileout
    integer :: thread_in_x, thread_in_y, thread_in_z
    integer :: thread_in_x_pascal, thread_in_y_pascal, thread_in_fft_pascal, Poisson_model
    character(len=128) :: dir_cont_filein, dir_cont_fileout, dir_instantfield

    namelist /domain_size/ L1, L2, L3
    namelist /mesh/ n1m, n2m, n3m
    namelist /grid_style/ pbc1, pbc2, pbc3, uniform1, uniform2, uniform3, gamma1, gamma2, gamma3
    namelist /MPI_procs/ np1, np2, np3
    namelist /cuda_thread/ thread_in_x, thread_in_y, thread_in_z
    namelist /cuda_pascal_tdma_thread/ thread_in_x_pascal, thread_in_y_pascal, thread_in_fft_pascal
    namelist /Poisson/ Poisson_model
    namelist /post_option/ ContinueFilein, ContinueFileout
    namelist /directories/ dir_cont_filein, dir_cont_fileout, dir_instantfield

    character(len=512) :: input
    integer :: u, ios, n1, n2, n3, i, j, k, sweeps
    real(rp), allocatable :: x1(:), x2(:), x3(:), dx1(:), dx2(:), dx3(:), dmx1(:), dmx2(:), dmx3(:)
    real(rp), allocatable :: prhs(:, :, :), rhs_dev(:, :, :), P(:, :, :), Pcont(:, :, :)
    real(rp), pointer :: PRHS_d(:, :, :) => null(), P_d(:, :, :) => null()
    real(rp), pointer :: dmx1_d(:) => null(), dmx2_d(:) => null(), dmx3_d(:) => null(), dx3_d(:) => null()
    real(rp), pointer :: x1_d(:) => null(), x2_d(:) => null(), x3_d(:) => null()
    integer(int64) :: c0, c1, rate
    real(rp) :: rel, rel_solver, rhs_diff

    ! defaults of the reference input file (01_Fortran/run/PARA_INPUT_GPU.dat)
    L1 = 2.0_rp; L2 = 2.0_rp; L3 = 2.0_rp
    n1m = 64; n2m = 64; n3m = 64
    pbc1 = .true.; pbc2 = .true.; pbc3 = .false.
    uniform1 = 1; uniform2 = 1; uniform3 = 1
    gamma1 = 3.0_rp; gamma2 = 3.0_rp; gamma3 = 3.0_rp
    np1 = 1; np2 = 1; np3 = 1
    thread_in_x = 8; thread_in_y = 4; thread_in_z = 4
    thread_in_x_pascal = 32; thread_in_y_pascal = 8; thread_in_fft_pascal = 32
    Poisson_model = 1
    ContinueFilein = .false.; ContinueFileout = .true.
    dir_cont_filein = './data/1_continue/MPI_IO/'
    dir_cont_fileout = './data/1_continue/'
    dir_instantfield = './data/2_instantfield/'

    input = '../run/PARA_INPUT.dat'
    if (command_argument_count() >= 1) call get_command_argument(1, input)
    open(newunit=u, file=trim(input), status='old', action='read', iostat=ios)
    if (ios /= 0) then
        write(*, '(a,a)') 'cannot open input file ', trim(input)
        error stop 1
    end if
    read(u, nml=domain_size)
    read(u, nml=mesh)
    read(u, nml=grid_style)
    read(u, nml=MPI_procs)
    read(u, nml=cuda_thread)
    read(u, nml=cuda_pascal_tdma_thread)
    read(u, nml=Poisson)
    read(u, nml=post_option)
    read(u, nml=directories)
    close(u)

    n1 = n1m + 1; n2 = n2m + 1; n3 = n3m + 1
    write(*, '(a)') '========== Poisson FFT solver (MI355X, one GPU) =========='
    write(*, '(a,3f10.2)') '  Domain  (L1,L2,L3)        :', L1, L2, L3
    write(*, '(a,3i10)')   '  Mesh    (n1m,n2m,n3m)     :', n1m, n2m, n3m
    write(*, '(a,3l10)')   '  PBC     (x1, x2, x3)      :', pbc1, pbc2, pbc3
    write(*, '(a,3i10)')   '  Uniform (x1, x2, x3)      :', uniform1, uniform2, uniform3
    if (np1 * np2 * np3 /= 1) write(*, '(a,3i4,a)') '  MPI_procs', np1, np2, np3, ' ignored: one process'
    if (mod(n1m, 2) /= 0 .or. mod(n2m, 2) /= 0 .or. mod(n3m, 2) /= 0) then
        write(*, '(a)') 'n1m, n2m, n3m must all be even. Aborting.'
        error stop 1
    end if

    ! grid metrics (mpi_subdomain.f90:397-472 on one process)
    allocate(x1(0:n1), x2(0:n2), x3(0:n3), dx1(0:n1), dx2(0:n2), dx3(0:n3), dmx1(0:n1), dmx2(0:n2), dmx3(0:n3))
    call make_axis(n1m, L1, pbc1, uniform1, gamma1, x1, dx1, dmx1)
    call make_axis(n2m, L2, pbc2, uniform2, gamma2, x2, dx2, dmx2)
    call make_axis(n3m, L3, pbc3, uniform3, gamma3, x3, dx3, dmx3)

    ! device arrays
    call ppf_device_alloc_3d(PRHS_d, [1, 1, 1], [n1m, n2m, n3m])
    call ppf_device_alloc_3d(P_d, [0, 0, 0], [n1, n2, n3])
    call ppf_device_alloc_1d(dmx1_d, 0, n1)
    call ppf_device_alloc_1d(dmx2_d, 0, n2)
    call ppf_device_alloc_1d(dmx3_d, 0, n3)
    call ppf_device_alloc_1d(dx3_d, 0, n3)
    call ppf_upload_1d(dmx1_d, dmx1)
    call ppf_upload_1d(dmx2_d, dmx2)
    call ppf_upload_1d(dmx3_d, dmx3)
    call ppf_upload_1d(dx3_d, dx3)
    ! node coordinates for the RHS kernel (cuda_subdomain.f90: x1_d, x2_d, x3_d)
    call ppf_device_alloc_1d(x1_d, 0, n1)
    call ppf_device_alloc_1d(x2_d, 0, n2)
    call ppf_device_alloc_1d(x3_d, 0, n3)
    call ppf_upload_1d(x1_d, x1)
    call ppf_upload_1d(x2_d, x2)
    call ppf_upload_1d(x3_d, x3)
    allocate(prhs(n1m, n2m, n3m), rhs_dev(n1m, n2m, n3m), P(0:n1, 0:n2, 0:n3))
    P = 0.0_rp
    call ppf_upload_3d(P_d, P)

    call fft_poisson_plan_cuda_create(0, 0, 0, 1, 1, 1, -1, -1, -1, -1, -1, -1, 0, 0, 0, &
                                      n1, n2, n3, n1, n2, n3, L1, L2, L3, pbc1, pbc2, pbc3, &
                                      dim3(thread_in_x_pascal, thread_in_y_pascal, 1), &
                                      dim3(thread_in_fft_pascal, thread_in_y_pascal, 1))
    call cuda_Poisson_FFT_initial()
    if (ContinueFilein) then
        ! cuda_Post_FileIn_Continue_Post_Reassembly_IO (cuda_post.f90:224-258) on
        ! one process: the interior from one direct-access record, periodic
        ! ghosts wrapped (mpi_subdomain_ghostcell_update_real), then P_d = P
        call read_continue_file(trim(dir_cont_filein)//'cont_P.bin', P)
        allocate(Pcont(n1m, n2m, n3m))
        Pcont = P(1:n1m, 1:n2m, 1:n3m)
        call ppf_upload_3d(P_d, P)
        write(*, '(a)') 'Read continue file using post reassembly IO'
    end if
    write(*, '(a)') '[[[============== all setup finished ==============]]]'

    ! timed as the reference times it (main.f90:44-61): RHS kernel, solve, D2H
    call system_clock(c0, rate)
    call cuda_pressure_RHS(PRHS_d)
    call cuda_Poisson_FFT_1D(PRHS_d, P_d, dmx1_d, dmx2_d, dmx3_d, dx3_d)
    call ppf_download_3d(P, P_d)
    call system_clock(c1)
    write(*, '(a,es14.6)') ' Total calculation time: ', real(c1 - c0, rp) / real(rate, rp)

    ! the device RHS against the same formula on the host (cuda_pressure.f90:87)
    call ppf_download_3d(rhs_dev, PRHS_d)
    do k = 1, n3m
        do j = 1, n2m
            do i = 1, n1m
                prhs(i, j, k) = -cos(x1(i) * PI) * cos(x2(j) * PI) * cos(x3(k) * PI) * 3.0_rp * PI * PI
            end do
        end do
    end do
    rhs_diff = maxval(abs(rhs_dev - prhs)) / maxval(abs(prhs))
    write(*, '(a,es14.6)') ' RHS device vs host (max rel diff): ', rhs_diff
    if (ContinueFilein) write(*, '(a,es14.6)') ' Continue-file field vs solution (max abs diff): ', &
        maxval(abs(Pcont - P(1:n1m, 1:n2m, 1:n3m)))

    rel = fft_poisson_residual(PRHS_d, P_d, dmx3_d, dx3_d)
    call fft_poisson_last_stats(sweeps, rel_solver)
    write(*, '(a,es14.6)') ' Relative residual: ', rel
    write(*, '(a,es24.16)') ' Interior sum of squares: ', sum(P(1:n1m, 1:n2m, 1:n3m) ** 2)

    if (ContinueFileout) then
        call execute_command_line('mkdir -p '//trim(dir_cont_fileout))
        open(newunit=u, file=trim(dir_cont_fileout)//'cont_P.bin', form='unformatted', access='stream', &
             status='replace', action='write')
        write(u) P(1:n1m, 1:n2m, 1:n3m)
        close(u)
        write(*, '(a)') 'Write continue file using post reassembly IO'
    end if

    call write_instant_field(trim(dir_instantfield), P, x1, x2, x3, dx1, dx2, dx3)

    call cuda_Poisson_FFT_clean()
    call ppf_device_free_3d(PRHS_d)
    call ppf_device_free_3d(P_d)
    call ppf_device_free_1d(dmx1_d)
    call ppf_device_free_1d(dmx2_d)
    call ppf_device_free_1d(dmx3_d)
    call ppf_device_free_1d(dx3_d)
    call ppf_device_free_1d(x1_d)
    call ppf_device_free_1d(x2_d)
    call ppf_device_free_1d(x3_d)
    write(*, '(a)') '[Main] Simulation complete. (GPU)'

contains

    !> The example's RHS on the device (cuda_pressure_RHS,
    !> cuda_pressure.f90:66-90): launched on the default stream, ordered before
    !> the solve.
    subroutine cuda_pressure_RHS(PRHS_d)
        real(rp), pointer, intent(in) :: PRHS_d(:, :, :)
        call ppf_check(c_example_rhs(c_loc(PRHS_d(1, 1, 1)), c_loc(x1_d(0)), c_loc(x2_d(0)), c_loc(x3_d(0)), &
                                     int(n1m, c_int), int(n2m, c_int), int(n3m, c_int), c_null_ptr), 'cuda_pressure_RHS')
    end subroutine cuda_pressure_RHS

    !> Interior P(1:n1m,1:n2m,1:n3m) from the continue file -- raw FP64 in
    !> Fortran order, one record -- and the one-process ghost exchange: wrap
    !> along periodic axes; wall ghosts keep their values.
    subroutine read_continue_file(path, P)
        character(len=*), intent(in) :: path
        real(rp), intent(inout) :: P(0:, 0:, 0:)
        real(rp), allocatable :: var_all(:, :, :)
        integer :: u, ios, nx, ny, nz
        nx = size(P, 1) - 2; ny = size(P, 2) - 2; nz = size(P, 3) - 2
        allocate(var_all(nx, ny, nz))
        open(newunit=u, file=path, form='unformatted', access='direct', recl=int(nx, int64) * ny * nz * &
             (storage_size(var_all) / 8), status='old', action='read', iostat=ios)
        if (ios /= 0) then
            write(*, '(a,a)') 'cannot open continue file ', path
            error stop 1
        end if
        read(u, rec=1) var_all
        close(u)
        P(1:nx, 1:ny, 1:nz) = var_all
        if (pbc1) then
            P(0, :, :) = P(nx, :, :); P(nx + 1, :, :) = P(1, :, :)
        end if
        if (pbc2) then
            P(:, 0, :) = P(:, ny, :); P(:, ny + 1, :) = P(:, 1, :)
        end if
        if (pbc3) then
            P(:, :, 0) = P(:, :, nz); P(:, :, nz + 1) = P(:, :, 1)
        end if
    end subroutine read_continue_file

    !> Tecplot ASCII output of the reference example (cuda_post.f90:294-331,
    !> calc_C2g :339-353): the value at node (x1(i), x2(j), x3(k)), i = 1..n1m
    !> etc., is the cell-centred P interpolated linearly in z, then x, then y
    !> over the 8 cells that share the node, each cell weighted by the width of
    !> the cell on the other side of the node.
    subroutine write_instant_field(dir, P, x1, x2, x3, dx1, dx2, dx3)
        character(len=*), intent(in) :: dir
        real(rp), intent(in) :: P(0:, 0:, 0:), x1(0:), x2(0:), x3(0:), dx1(0:), dx2(0:), dx3(0:)
        integer :: u, i, j, k, nx, ny, nz
        real(rp) :: wz0, wz1, wx0, wx1, wy0, wy1, lo, hi
        nx = size(P, 1) - 2; ny = size(P, 2) - 2; nz = size(P, 3) - 2
        call execute_command_line('mkdir -p '//dir)
        open(newunit=u, file=dir//'P_000.PLT', status='replace', action='write')
        write(u, '(a)') ' VARIABLES="X","Y","Z","P"'
        write(u, '(a,i0,a,i0,a,i0,a,i0)') ' zone t="', 1, '" i=', nx, ' j=', ny, ' k=', nz
        do k = 1, nz
            ! weight of plane k is the width of plane k-1, and vice versa
            wz1 = dx3(k - 1) / (dx3(k) + dx3(k - 1)); wz0 = dx3(k) / (dx3(k) + dx3(k - 1))
            do j = 1, ny
                wy1 = dx2(j - 1) / (dx2(j) + dx2(j - 1)); wy0 = dx2(j) / (dx2(j) + dx2(j - 1))
                do i = 1, nx
                    wx1 = dx1(i - 1) / (dx1(i) + dx1(i - 1)); wx0 = dx1(i) / (dx1(i) + dx1(i - 1))
                    lo = wx1 * (wz1 * P(i, j - 1, k) + wz0 * P(i, j - 1, k - 1)) &
                       + wx0 * (wz1 * P(i - 1, j - 1, k) + wz0 * P(i - 1, j - 1, k - 1))
                    hi = wx1 * (wz1 * P(i, j, k) + wz0 * P(i, j, k - 1)) &
                       + wx0 * (wz1 * P(i - 1, j, k) + wz0 * P(i - 1, j, k - 1))
                    write(u, '(4(E11.4,1x))') x1(i), x2(j), x3(k), wy1 * hi + wy0 * lo
                end do
            end do
        end do
        close(u)
    end subroutine write_instant_field

    !> Nodes, widths and centre spacings of one axis on one process:
    !> calculate_x / uniform_grid / non_uniform_channel_grid /
    !> calculate_dx_dmx with the one-process ghost exchange.
    subroutine make_axis(nm, L, pbc, uniform, gamma, x, dx, dmx)
        integer, intent(in) :: nm, uniform
        real(rp), intent(in) :: L, gamma
        logical, intent(in) :: pbc
        real(rp), intent(out) :: x(0:), dx(0:), dmx(0:)
        integer :: q, n
        n = nm + 1
        do q = 0, n
            if (uniform == 1) then
                x(q) = real(q - 1, rp) * L / real(nm, rp)
            else
                x(q) = L * 0.5_rp * (1.0_rp + tanh(0.5_rp * gamma * (2.0_rp * real(q - 1, rp) / real(nm, rp) - 1.0_rp)) &
                                      / tanh(gamma * 0.5_rp))
            end if
        end do
        if (.not. pbc) x(0) = x(1)
        dx = 0.0_rp
        dmx = 0.0_rp
        do q = 1, nm
            dx(q) = x(q + 1) - x(q)
        end do
        if (pbc) then
            dx(0) = dx(nm); dx(n) = dx(1)
        end if
        do q = 1, nm
            dmx(q) = 0.5_rp * (dx(q - 1) + dx(q))
        end do
        if (pbc) then
            dmx(0) = dmx(nm); dmx(n) = dmx(1)
        else
            dmx(n) = 0.5_rp * (dx(nm) + dx(n))
        end if
    end subroutine make_axis

end program ppf_example
