!> Fortran interface of the MI355X Poisson solve path.
!>
!> Keeps the procedure names and argument lists of the reference module
!> `fft_poisson` (01_Fortran/src/fft_3d_poisson.f90:1-130,353-361) so a caller
!> written against it compiles unchanged:
!>
!>   call fft_poisson_plan_cuda_create(rank1, ..., threads_tdma, threads_fft)
!>   call cuda_Poisson_FFT_initial()
!>   call cuda_Poisson_FFT_1D(PRHS_d, P_d, dmx1_d, dmx2_d, dmx3_d, dx3_d)
!>   call cuda_Poisson_FFT_clean()
!>
!> Every call goes to the C ABI of libpascal_poisson_fft (HIP, gfx950) through
!> ISO_C_BINDING.  "Device arrays" are Fortran pointers over ppf_malloc memory
!> (see ppf_device_alloc_* below); the host never dereferences them.
!>
!> Differences a caller can observe:
!>   * one GPU, one rank: the communicator arguments are accepted and must
!>     describe a single process along every axis;
!>   * the thread-shape arguments are accepted and ignored (launch shapes are
!>     chosen by the library per extent);
!>   * PRHS_d is left unmodified (the reference uses it as scratch);
!>   * errors print the library's message and stop, as the reference's own
!>     plan checks do (pascal_tdma_cuda.f90:168-190).
module fft_poisson
    use iso_c_binding
    implicit none

    private

    integer, parameter, public :: rp = kind(0.0d0)
    real(rp), parameter, public :: PI = real(acos(-1.0d0), rp)

    !> Launch-shape triple with the layout of CUDA Fortran's dim3.
    type, bind(C), public :: dim3
        integer(c_int) :: x, y, z
    end type dim3

    !> Plan scalars (fft_3d_poisson.f90:26-41).
    type, public :: fft_poisson_plan_cuda
        private
        integer  :: n1 = 0, n2 = 0, n3 = 0
        integer  :: n1sub = 0, n2sub = 0, n3sub = 0
        integer  :: n1m = 0, n2m = 0, n3m = 0
        integer  :: nprocs(3) = 1
        real(rp) :: L1 = 0, L2 = 0, L3 = 0
        logical  :: pbc1 = .false., pbc2 = .false., pbc3 = .false.
        type(dim3) :: threads_tdma, threads_fft
    end type fft_poisson_plan_cuda

    integer, parameter, public :: PPF_PREC_FP64 = 0, PPF_PREC_FP32_IR = 1

    type(fft_poisson_plan_cuda), public :: p_poi

    public :: fft_poisson_plan_cuda_create, cuda_Poisson_FFT_initial, cuda_Poisson_FFT_clean, cuda_Poisson_FFT_1D
    public :: fft_poisson_set_precision, fft_poisson_last_stats, fft_poisson_residual
    public :: ppf_device_alloc_1d, ppf_device_alloc_3d, ppf_device_free_1d, ppf_device_free_3d
    public :: ppf_upload_1d, ppf_upload_3d, ppf_download_3d, ppf_check

    ! ---- C ABI (include/pascal_poisson_fft.h) --------------------------------
    type, bind(C) :: ppf_desc_t
        integer(c_int) :: n1m, n2m, n3m
        real(c_double) :: L1, L2, L3
        integer(c_int) :: pbc1, pbc2, pbc3
        integer(c_int) :: precision
        real(c_double) :: tol
        integer(c_int) :: max_sweeps
    end type ppf_desc_t

    type, bind(C) :: ppf_stats_t
        integer(c_int) :: sweeps
        real(c_double) :: rel_residual
    end type ppf_stats_t

    interface
        integer(c_int) function c_plan_create(desc, plan) bind(C, name='ppf_plan_create')
            import :: c_int, c_ptr, ppf_desc_t
            type(ppf_desc_t), intent(in) :: desc
            type(c_ptr), intent(out) :: plan
        end function
        integer(c_int) function c_solve(plan, prhs, p, dmx1, dmx2, dmx3, dx3, stream, stats) &
                bind(C, name='ppf_solve')
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, prhs, p, dmx1, dmx2, dmx3, dx3, stream, stats
        end function
        integer(c_int) function c_residual(plan, f, p, dmx3, dx3, stream, rel) bind(C, name='ppf_residual')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, f, p, dmx3, dx3, stream
            real(c_double), intent(out) :: rel
        end function
        integer(c_int) function c_plan_destroy(plan) bind(C, name='ppf_plan_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
        end function
        type(c_ptr) function c_error_string(code) bind(C, name='ppf_error_string')
            import :: c_ptr, c_int
            integer(c_int), value :: code
        end function
        integer(c_int) function c_malloc(ptr, bytes) bind(C, name='ppf_malloc')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), intent(out) :: ptr
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function c_free(ptr) bind(C, name='ppf_free')
            import :: c_int, c_ptr
            type(c_ptr), value :: ptr
        end function
        integer(c_int) function c_memcpy(dst, src, bytes, kind) bind(C, name='ppf_memcpy')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
            integer(c_int), value :: kind
        end function
        integer(c_int) function c_sync(stream) bind(C, name='ppf_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: stream
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    integer(c_int), parameter :: H2D = 1, D2H = 2

    type(c_ptr), save :: plan_handle = c_null_ptr
    logical, save :: created = .false.
    integer, save :: prec_mode = PPF_PREC_FP64, prec_sweeps = 0
    real(rp), save :: prec_tol = 0.0_rp
    type(ppf_stats_t), target, save :: last_stats = ppf_stats_t(0, -1.0d0)

contains

    !> Stops with the library's message when rc is not PPF_OK.
    subroutine ppf_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: s
        integer :: n
        if (rc == 0) return
        s = c_error_string(rc)
        n = int(c_strlen(s))
        call c_f_pointer(s, msg, [n])
        write(*, '(a,a,a,i0,a,*(a))') '[fft_poisson] ', what, ' failed (', rc, '): ', msg
        error stop 1
    end subroutine ppf_check

    !> fft_3d_poisson.f90:70-108.  n = nm + 1 per axis (:103-104).
    subroutine fft_poisson_plan_cuda_create(rank1, rank2, rank3, np1, np2, np3, wrank1, wrank2, wrank3, &
                                            erank1, erank2, erank3, comm1, comm2, comm3, &
                                            n1, n2, n3, n1sub, n2sub, n3sub, L1, L2, L3, pbc1, pbc2, pbc3, &
                                            threads_tdma, threads_fft)
        integer, intent(in)    :: rank1, rank2, rank3
        integer, intent(in)    :: np1, np2, np3
        integer, intent(in)    :: wrank1, wrank2, wrank3
        integer, intent(in)    :: erank1, erank2, erank3
        integer, intent(in)    :: comm1, comm2, comm3
        integer, intent(in)    :: n1, n2, n3
        integer, intent(in)    :: n1sub, n2sub, n3sub
        real(rp), intent(in)   :: L1, L2, L3
        logical, intent(in)    :: pbc1, pbc2, pbc3
        type(dim3), intent(in) :: threads_tdma, threads_fft

        ! ranks, neighbour ranks and communicators carry no information for one
        ! process; the process counts are checked by cuda_Poisson_FFT_initial
        p_poi%n1 = n1; p_poi%n2 = n2; p_poi%n3 = n3
        p_poi%n1sub = n1sub; p_poi%n2sub = n2sub; p_poi%n3sub = n3sub
        p_poi%n1m = n1 - 1; p_poi%n2m = n2 - 1; p_poi%n3m = n3 - 1
        p_poi%nprocs = [np1, np2, np3]
        p_poi%L1 = L1; p_poi%L2 = L2; p_poi%L3 = L3
        p_poi%pbc1 = pbc1; p_poi%pbc2 = pbc2; p_poi%pbc3 = pbc3
        p_poi%threads_tdma = threads_tdma; p_poi%threads_fft = threads_fft
    end subroutine fft_poisson_plan_cuda_create

    !> Extension: solve precision for the next cuda_Poisson_FFT_initial.
    !> prec = PPF_PREC_FP64 (default) or PPF_PREC_FP32_IR with its relative
    !> residual target tol and sweep cap (0 selects the library defaults).
    subroutine fft_poisson_set_precision(prec, tol, max_sweeps)
        integer, intent(in) :: prec
        real(rp), intent(in), optional :: tol
        integer, intent(in), optional :: max_sweeps
        prec_mode = prec
        prec_tol = 0.0_rp
        prec_sweeps = 0
        if (present(tol)) prec_tol = tol
        if (present(max_sweeps)) prec_sweeps = max_sweeps
    end subroutine fft_poisson_set_precision

    !> fft_3d_poisson.f90:110-121: BC type, buffers, transform tables and
    !> Thomas coefficients -- one ppf_plan_create.
    subroutine cuda_Poisson_FFT_initial()
        type(ppf_desc_t) :: d
        if (any(p_poi%nprocs /= 1)) then
            write(*, '(a,3i4,a)') '[fft_poisson] process grid', p_poi%nprocs, &
                ' requested; this library solves on one GPU (one process per axis)'
            error stop 1
        end if
        if (p_poi%n1sub /= p_poi%n1 .or. p_poi%n2sub /= p_poi%n2 .or. p_poi%n3sub /= p_poi%n3) then
            write(*, '(a)') '[fft_poisson] n1sub/n2sub/n3sub must equal n1/n2/n3 on one process'
            error stop 1
        end if
        if (created) call cuda_Poisson_FFT_clean()
        d%n1m = p_poi%n1m; d%n2m = p_poi%n2m; d%n3m = p_poi%n3m
        d%L1 = p_poi%L1; d%L2 = p_poi%L2; d%L3 = p_poi%L3
        d%pbc1 = merge(1, 0, p_poi%pbc1); d%pbc2 = merge(1, 0, p_poi%pbc2); d%pbc3 = merge(1, 0, p_poi%pbc3)
        d%precision = prec_mode
        d%tol = prec_tol
        d%max_sweeps = prec_sweeps
        call ppf_check(c_plan_create(d, plan_handle), 'cuda_Poisson_FFT_initial')
        created = .true.
    end subroutine cuda_Poisson_FFT_initial

    !> fft_3d_poisson.f90:123-130.
    subroutine cuda_Poisson_FFT_clean()
        if (.not. created) return
        call ppf_check(c_plan_destroy(plan_handle), 'cuda_Poisson_FFT_clean')
        plan_handle = c_null_ptr
        created = .false.
    end subroutine cuda_Poisson_FFT_clean

    !> fft_3d_poisson.f90:353-626.  PRHS_d(1:n1m,1:n2m,1:n3m) is read;
    !> P_d(0:n1m+1,0:n2m+1,0:n3m+1) receives the zero-mean solution with its
    !> wall and periodic ghost layers; the metrics hold the reference's
    !> (0:nsub) arrays.  Returns after the device work has finished.
    subroutine cuda_Poisson_FFT_1D(PRHS_d, P_d, dmx1_d, dmx2_d, dmx3_d, dx3_d)
        real(rp), dimension(1:, 1:, 1:), target :: PRHS_d
        real(rp), dimension(0:, 0:, 0:), target :: P_d
        real(rp), dimension(:), target :: dmx1_d, dmx2_d, dmx3_d
        real(rp), dimension(:), target :: dx3_d

        if (.not. created) then
            write(*, '(a)') '[fft_poisson] cuda_Poisson_FFT_1D called before cuda_Poisson_FFT_initial'
            error stop 1
        end if
        if (size(PRHS_d, 1) /= p_poi%n1m .or. size(PRHS_d, 2) /= p_poi%n2m .or. size(PRHS_d, 3) /= p_poi%n3m .or. &
            size(P_d, 1) /= p_poi%n1m + 2 .or. size(P_d, 2) /= p_poi%n2m + 2 .or. size(P_d, 3) /= p_poi%n3m + 2 .or. &
            size(dmx1_d) < p_poi%n1m + 2 .or. size(dmx2_d) < p_poi%n2m + 2 .or. &
            size(dmx3_d) < p_poi%n3m + 2 .or. size(dx3_d) < p_poi%n3m + 2) then
            write(*, '(a)') '[fft_poisson] cuda_Poisson_FFT_1D: array extents do not match the plan'
            error stop 1
        end if
        if (.not. (is_contiguous(PRHS_d) .and. is_contiguous(P_d) .and. is_contiguous(dmx1_d) .and. &
                   is_contiguous(dmx2_d) .and. is_contiguous(dmx3_d) .and. is_contiguous(dx3_d))) then
            write(*, '(a)') '[fft_poisson] cuda_Poisson_FFT_1D: arrays must be contiguous device allocations'
            error stop 1
        end if
        call ppf_check(c_solve(plan_handle, c_loc(PRHS_d(1, 1, 1)), c_loc(P_d(0, 0, 0)), c_loc(dmx1_d(1)), &
                               c_loc(dmx2_d(1)), c_loc(dmx3_d(1)), c_loc(dx3_d(1)), c_null_ptr, c_loc(last_stats)), &
                       'cuda_Poisson_FFT_1D')
        call ppf_check(c_sync(c_null_ptr), 'cuda_Poisson_FFT_1D (sync)')
    end subroutine cuda_Poisson_FFT_1D

    !> Extension: refinement sweeps and final relative residual of the last
    !> solve (sweeps = 0 and residual < 0 for FP64 plans).
    subroutine fft_poisson_last_stats(sweeps, rel_residual)
        integer, intent(out) :: sweeps
        real(rp), intent(out) :: rel_residual
        sweeps = last_stats%sweeps
        rel_residual = last_stats%rel_residual
    end subroutine fft_poisson_last_stats

    !> Extension: FP64 relative residual ||f - L p|| / ||f|| of the operator
    !> the solve inverts, for device arrays shaped as in cuda_Poisson_FFT_1D.
    real(rp) function fft_poisson_residual(F_d, P_d, dmx3_d, dx3_d) result(rel)
        real(rp), dimension(1:, 1:, 1:), target :: F_d
        real(rp), dimension(0:, 0:, 0:), target :: P_d
        real(rp), dimension(:), target :: dmx3_d, dx3_d
        real(c_double) :: r
        call ppf_check(c_residual(plan_handle, c_loc(F_d(1, 1, 1)), c_loc(P_d(0, 0, 0)), c_loc(dmx3_d(1)), &
                                  c_loc(dx3_d(1)), c_null_ptr, r), 'fft_poisson_residual')
        rel = r
    end function fft_poisson_residual

    ! ---- device arrays ---------------------------------------------------------
    ! Stand-ins for CUDA Fortran `device, allocatable` arrays
    ! (01_Fortran/examples/cuda_subdomain.f90:96-180): pointers with the
    ! requested bounds over device memory.

    subroutine ppf_device_alloc_1d(a, lo, hi)
        real(rp), pointer, intent(out) :: a(:)
        integer, intent(in) :: lo, hi
        real(rp), pointer :: flat(:)
        type(c_ptr) :: cp
        integer(c_size_t) :: n
        n = int(max(hi - lo + 1, 1), c_size_t)
        call ppf_check(c_malloc(cp, n * 8_c_size_t), 'ppf_device_alloc_1d')
        call c_f_pointer(cp, flat, [n])
        a(lo:hi) => flat
    end subroutine ppf_device_alloc_1d

    subroutine ppf_device_alloc_3d(a, lo, hi)
        real(rp), pointer, intent(out) :: a(:, :, :)
        integer, intent(in) :: lo(3), hi(3)
        real(rp), pointer :: flat(:)
        type(c_ptr) :: cp
        integer(c_size_t) :: n
        n = int(hi(1) - lo(1) + 1, c_size_t) * int(hi(2) - lo(2) + 1, c_size_t) * int(hi(3) - lo(3) + 1, c_size_t)
        call ppf_check(c_malloc(cp, n * 8_c_size_t), 'ppf_device_alloc_3d')
        call c_f_pointer(cp, flat, [n])
        a(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)) => flat
    end subroutine ppf_device_alloc_3d

    subroutine ppf_device_free_1d(a)
        real(rp), pointer, intent(inout) :: a(:)
        if (.not. associated(a)) return
        call ppf_check(c_free(c_loc(a(lbound(a, 1)))), 'ppf_device_free_1d')
        nullify(a)
    end subroutine ppf_device_free_1d

    subroutine ppf_device_free_3d(a)
        real(rp), pointer, intent(inout) :: a(:, :, :)
        if (.not. associated(a)) return
        call ppf_check(c_free(c_loc(a(lbound(a, 1), lbound(a, 2), lbound(a, 3)))), 'ppf_device_free_3d')
        nullify(a)
    end subroutine ppf_device_free_3d

    !> Host -> device copy of a whole 1-D array (d = h).
    subroutine ppf_upload_1d(d, h)
        real(rp), pointer, intent(in) :: d(:)
        real(rp), target, contiguous, intent(in) :: h(:)
        call ppf_check(c_memcpy(c_loc(d(lbound(d, 1))), c_loc(h), int(size(h), c_size_t) * 8_c_size_t, H2D), &
                       'ppf_upload_1d')
    end subroutine ppf_upload_1d

    !> Host -> device copy of a whole 3-D array (d = h).
    subroutine ppf_upload_3d(d, h)
        real(rp), pointer, intent(in) :: d(:, :, :)
        real(rp), target, contiguous, intent(in) :: h(:, :, :)
        call ppf_check(c_memcpy(c_loc(d(lbound(d, 1), lbound(d, 2), lbound(d, 3))), c_loc(h), &
                                int(size(h, kind=c_size_t), c_size_t) * 8_c_size_t, H2D), 'ppf_upload_3d')
    end subroutine ppf_upload_3d

    !> Device -> host copy of a whole 3-D array (h = d), the counterpart of
    !> cuda_subdomain_DtoH.
    subroutine ppf_download_3d(h, d)
        real(rp), target, contiguous, intent(inout) :: h(:, :, :)
        real(rp), pointer, intent(in) :: d(:, :, :)
        call ppf_check(c_memcpy(c_loc(h), c_loc(d(lbound(d, 1), lbound(d, 2), lbound(d, 3))), &
                                int(size(h, kind=c_size_t), c_size_t) * 8_c_size_t, D2H), 'ppf_download_3d')
    end subroutine ppf_download_3d

end module fft_poisson
