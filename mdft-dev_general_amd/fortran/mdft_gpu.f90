! mdft_gpu.f90 -- Fortran host binding of libmdft_gpu.so (include/mdft_gpu.h).
!
! Mirrors the reference's evaluation interface:
!   energy_and_gradient(f, df)   src/module_energy_and_gradient.f90:36,63-64
!   excess_energy(ff, df, print_timers)
!                                src/module_proc_ptrs.f90:3-21 (excess_energy_t)
! The unknown xi(no,nx,ny,nz) lives on the device (xi_dev); the caller keeps a
! host copy only when it needs one.  A non-zero status from the library is
! turned into `error stop` with the library's message, as the reference stops
! on the same conditions.
module mdft_gpu
    use iso_c_binding
    implicit none
    private

    integer, parameter :: dp = c_double
    integer(c_int), parameter, public :: MDFT_PREC_FP64 = 0, MDFT_PREC_FP32 = 1, MDFT_PREC_BF16 = 2
    integer(c_int), parameter, public :: MDFT_DEVICE_PTRS = 1, MDFT_TIMERS = 2, MDFT_DF_FP32 = 4, MDFT_SCALED_VARS = 8
    integer(c_int), parameter, public :: MDFT_XI_FP32 = 16

    type, bind(C), public :: mdft_desc
        integer(c_int) :: nx, ny, nz
        real(c_double) :: lx, ly, lz
        integer(c_int) :: mmax, mrso
        real(c_double) :: kT, n0
        integer(c_int) :: precision, device
    end type

    ! ---- C-ABI
    interface
        integer(c_int) function mdft_ctx_create(desc, ctx) bind(C, name="mdft_ctx_create")
            import :: c_int, c_ptr, mdft_desc
            type(mdft_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine mdft_ctx_destroy(ctx) bind(C, name="mdft_ctx_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        type(c_ptr) function mdft_last_error(ctx) bind(C, name="mdft_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mdft_get_info(ctx, info) bind(C, name="mdft_get_info")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: info(10)
        end function
        integer(c_int) function mdft_set_ck(ctx, np_file, m, n, mu, nu, khi, nq_file, dq, c) bind(C, name="mdft_set_ck")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: np_file, nq_file
            integer(c_int), intent(in) :: m(*), n(*), mu(*), nu(*), khi(*)
            real(c_double), value :: dq
            real(c_double), intent(in) :: c(*)
        end function
        integer(c_int) function mdft_compute_vext(ctx, natom, atoms, nsite, sites, rcut) bind(C, name="mdft_compute_vext")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: natom, nsite
            real(c_double), intent(in) :: atoms(*), sites(*)
            real(c_double), value :: rcut
        end function
        ! electrostatics: MDFT_ELEC_POISSON (0, the reference's fast Poisson solver),
        ! MDFT_ELEC_DIRECT (1, minimum-image Coulomb) or MDFT_ELEC_NONE (2)
        integer(c_int) function mdft_compute_vext_ex(ctx, natom, atoms, nsite, sites, rcut, elec) &
                bind(C, name="mdft_compute_vext_ex")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: natom, nsite, elec
            real(c_double), intent(in) :: atoms(*), sites(*)
            real(c_double), value :: rcut
        end function
        integer(c_int) function mdft_set_vext(ctx, vext, flags) bind(C, name="mdft_set_vext")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, vext
            integer(c_int), value :: flags
        end function
        integer(c_int) function mdft_energy_and_gradient(ctx, xi, f_parts, df, flags) bind(C, name="mdft_energy_and_gradient")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, xi, df
            real(c_double), intent(out) :: f_parts(3)
            integer(c_int), value :: flags
        end function
        ! the same, with dots(j) = df . vecs(j), j <= k <= 3, formed in the final pass
        integer(c_int) function mdft_energy_and_gradient_dots(ctx, xi, f_parts, df, flags, k, vecs, dots) &
                bind(C, name="mdft_energy_and_gradient_dots")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, xi, df
            real(c_double), intent(out) :: f_parts(3)
            integer(c_int), value :: flags, k
            type(c_ptr), intent(in) :: vecs(*)
            real(c_double), intent(out) :: dots(*)
        end function
        integer(c_int) function mdft_excess_energy(ctx, xi, fexc, df, flags) bind(C, name="mdft_excess_energy")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, xi, df
            real(c_double), intent(out) :: fexc
            integer(c_int), value :: flags
        end function
        integer(c_int) function mdft_device_alloc(ctx, bytes, dptr) bind(C, name="mdft_device_alloc")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx
            integer(c_size_t), value :: bytes
            type(c_ptr), intent(out) :: dptr
        end function
        integer(c_int) function mdft_device_free(ctx, dptr) bind(C, name="mdft_device_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, dptr
        end function
        integer(c_int) function mdft_memcpy_h2d(ctx, dst, src, bytes) bind(C, name="mdft_memcpy_h2d")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, dst, src
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function mdft_memcpy_d2h(ctx, dst, src, bytes) bind(C, name="mdft_memcpy_d2h")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, dst, src
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function mdft_vec_dot(ctx, n, x, y, out) bind(C, name="mdft_vec_dot")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, x, y
            integer(c_size_t), value :: n
            real(c_double), intent(out) :: out
        end function
        integer(c_int) function mdft_vec_axpy_dot(ctx, n, a, x, y, z, out) bind(C, name="mdft_vec_axpy_dot")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, x, y, z
            integer(c_size_t), value :: n
            real(c_double), value :: a
            real(c_double), intent(out) :: out
        end function
        integer(c_int) function mdft_vec_lincomb(ctx, n, a, x, b, y, z) bind(C, name="mdft_vec_lincomb")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, x, y, z
            integer(c_size_t), value :: n
            real(c_double), value :: a, b
        end function
        integer(c_int) function mdft_vec_multi_dot(ctx, n, k, vecs, x, out) bind(C, name="mdft_vec_multi_dot")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, x
            integer(c_size_t), value :: n
            integer(c_int), value :: k
            type(c_ptr), intent(in) :: vecs(*)
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function mdft_vec_combine(ctx, n, k, coef, vecs, z) bind(C, name="mdft_vec_combine")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, z
            integer(c_size_t), value :: n
            integer(c_int), value :: k
            real(c_double), intent(in) :: coef(*)
            type(c_ptr), intent(in) :: vecs(*)
        end function
        ! z = sum coef(j) vecs(j) and y = x + a z in one pass
        integer(c_int) function mdft_vec_combine_step(ctx, n, k, coef, vecs, z, a, x, y) &
                bind(C, name="mdft_vec_combine_step")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, z, x, y
            integer(c_size_t), value :: n
            integer(c_int), value :: k
            real(c_double), intent(in) :: coef(*)
            type(c_ptr), intent(in) :: vecs(*)
            real(c_double), value :: a
        end function
        ! the same passes over FP32-stored directions and gradients (FP64 iterate x, y)
        integer(c_int) function mdft_vec_multi_dot_f32(ctx, n, k, vecs, x, out) bind(C, name="mdft_vec_multi_dot_f32")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, x
            integer(c_size_t), value :: n
            integer(c_int), value :: k
            type(c_ptr), intent(in) :: vecs(*)
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function mdft_vec_combine_step_f32(ctx, n, k, coef, vecs, z, a, x, y) &
                bind(C, name="mdft_vec_combine_step_f32")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, z, x, y
            integer(c_size_t), value :: n
            integer(c_int), value :: k
            real(c_double), intent(in) :: coef(*)
            type(c_ptr), intent(in) :: vecs(*)
            real(c_double), value :: a
        end function
        integer(c_int) function mdft_vec_axpy_f32(ctx, n, a, z, x, y) bind(C, name="mdft_vec_axpy_f32")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, z, x, y
            integer(c_size_t), value :: n
            real(c_double), value :: a
        end function
        integer(c_int) function mdft_vec_widen_f32(ctx, n, in, out) bind(C, name="mdft_vec_widen_f32")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, in, out
            integer(c_size_t), value :: n
        end function
        integer(c_int) function mdft_vec_narrow_f64(ctx, n, in, out) bind(C, name="mdft_vec_narrow_f64")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, in, out
            integer(c_size_t), value :: n
        end function
        integer(c_int) function mdft_vec_axpy_f32x(ctx, n, a, z, x, y) bind(C, name="mdft_vec_axpy_f32x")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, z, x, y
            integer(c_size_t), value :: n
            real(c_double), value :: a
        end function
        integer(c_int) function mdft_vec_combine_step_f32x(ctx, n, k, coef, vecs, z, a, x, y) &
                bind(C, name="mdft_vec_combine_step_f32x")
            import :: c_int, c_ptr, c_size_t, c_double
            type(c_ptr), value :: ctx, z, x, y
            integer(c_size_t), value :: n
            integer(c_int), value :: k
            real(c_double), intent(in) :: coef(*)
            type(c_ptr), intent(in) :: vecs(*)
            real(c_double), value :: a
        end function
        ! u = s xi (dir = 1) or xi = u / s (dir = -1), s(io) = sqrt(w(io) / mean w)
        integer(c_int) function mdft_scale_variables(ctx, n, dir, in_fp32, in, out) bind(C, name="mdft_scale_variables")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx, in, out
            integer(c_size_t), value :: n
            integer(c_int), value :: dir, in_fp32
        end function
        integer(c_int) function mdft_charged_solute_corrections(lengths, n0, natom, atoms, nsite, sites, eps_r, out) &
                bind(C, name="mdft_charged_solute_corrections")
            import :: c_int, c_double
            real(c_double), intent(in) :: lengths(3), atoms(6, *), sites(6, *)
            real(c_double), value :: n0, eps_r
            integer(c_int), value :: natom, nsite
            real(c_double), intent(out) :: out(2)
        end function
    end interface

    public :: mdft_ctx_create, mdft_ctx_destroy, mdft_get_info, mdft_set_ck, mdft_compute_vext, mdft_compute_vext_ex, mdft_set_vext
    public :: mdft_energy_and_gradient, mdft_energy_and_gradient_dots, mdft_excess_energy, mdft_device_alloc
    public :: mdft_device_free
    public :: mdft_memcpy_h2d, mdft_memcpy_d2h, mdft_vec_dot, mdft_vec_axpy_dot, mdft_vec_lincomb
    public :: mdft_vec_multi_dot, mdft_vec_combine, mdft_vec_combine_step
    public :: mdft_vec_multi_dot_f32, mdft_vec_combine_step_f32, mdft_vec_axpy_f32, mdft_vec_widen_f32
    public :: mdft_vec_narrow_f64, mdft_vec_axpy_f32x, mdft_vec_combine_step_f32x
    public :: mdft_scale_variables
    public :: mdft_charged_solute_corrections
    public :: check, last_error_text

    ! ---- module state, the analogue of the reference's solvent(1)%xi and grid globals
    type(c_ptr), public :: ctx = c_null_ptr
    type(c_ptr), public :: xi_dev = c_null_ptr      ! xi(no,nx,ny,nz) on the device
    type(c_ptr), public :: df_dev = c_null_ptr      ! scratch gradient on the device
    integer(c_size_t), public :: nvar = 0           ! no*nx*ny*nz
    real(dp), public :: f_id = 0, f_ext = 0, f_exc = 0, f_tot = 0
    ! Hunenberger's corrections for a charged solute (module_energy_and_gradient.f90:21-26,190-203):
    ! set by set_charged_solute_corrections, added to f while the flag is set
    real(dp), public :: pbc_correction = 0, pscheme_correction = 0
    logical, public :: apply_energy_corrections_due_to_charged_solute = .true.

    public :: gpu_init, gpu_finalize, energy_and_gradient, excess_energy_gpu, set_charged_solute_corrections

contains

    function last_error_text(h) result(txt)
        type(c_ptr), intent(in) :: h
        character(len=:), allocatable :: txt
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: i
        p = mdft_last_error(h)
        call c_f_pointer(p, s, [4096])
        i = 0
        do while (i < 4096)
            if (s(i + 1) == c_null_char) exit
            i = i + 1
        end do
        allocate(character(len=i) :: txt)
        txt = transfer(s(1:i), txt)
    end function

    subroutine check(rc, where)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: where
        if (rc /= 0) then
            print '(3a,i0,2a)', "mdft_gpu: ", where, " failed (", rc, "): ", last_error_text(ctx)
            error stop "mdft_gpu: library error"
        end if
    end subroutine

    ! Create the device context and the device copy of xi (initialised to 1).
    subroutine gpu_init(desc)
        type(mdft_desc), intent(in) :: desc
        integer(c_int) :: info(10)
        integer(c_int) :: rc
        real(dp), allocatable, target :: ones(:)
        rc = mdft_ctx_create(desc, ctx)
        if (rc /= 0) then
            print '(2a)', "mdft_gpu: mdft_ctx_create failed: ", last_error_text(c_null_ptr)
            error stop "mdft_gpu: cannot create the device context"
        end if
        call check(mdft_get_info(ctx, info), "get_info")
        nvar = int(info(4), c_size_t) * int(desc%nx, c_size_t) * int(desc%ny, c_size_t) * int(desc%nz, c_size_t)
        call check(mdft_device_alloc(ctx, 8_c_size_t * nvar, xi_dev), "alloc xi")
        call check(mdft_device_alloc(ctx, 8_c_size_t * nvar, df_dev), "alloc df")
        allocate(ones(nvar), source=1.0_dp)
        call check(mdft_memcpy_h2d(ctx, xi_dev, c_loc(ones), 8_c_size_t * nvar), "upload xi")
    end subroutine

    subroutine gpu_finalize()
        if (c_associated(xi_dev)) call check(mdft_device_free(ctx, xi_dev), "free xi")
        if (c_associated(df_dev)) call check(mdft_device_free(ctx, df_dev), "free df")
        if (c_associated(ctx)) call mdft_ctx_destroy(ctx)
        ctx = c_null_ptr
        xi_dev = c_null_ptr
        df_dev = c_null_ptr
    end subroutine

    ! Same signature and meaning as the reference (module_energy_and_gradient.f90:36):
    ! f = F_id + F_ext + F_exc of the current xi; df (optional) = (1/dv) dF/dxi.
    subroutine energy_and_gradient(f, df)
        real(dp), intent(out) :: f
        real(dp), intent(out), optional, contiguous :: df(:,:,:,:,:)
        real(dp) :: parts(3)
        if (present(df)) then
            call check(mdft_energy_and_gradient(ctx, xi_dev, parts, df_dev, MDFT_DEVICE_PTRS), "energy_and_gradient")
            call from_device(df_dev, df)
        else
            call check(mdft_energy_and_gradient(ctx, xi_dev, parts, c_null_ptr, MDFT_DEVICE_PTRS), "energy_and_gradient")
        end if
        f_id = parts(1)
        f_ext = parts(2)
        f_exc = parts(3)
        f = parts(1) + parts(2) + parts(3)
        if (apply_energy_corrections_due_to_charged_solute) f = f + pbc_correction + pscheme_correction
        f_tot = f
    end subroutine

    ! Copies between a caller's array of nvar values and the device.  The caller's dummy
    ! arguments keep the reference's attributes (no TARGET: excess_energy_gpu must match
    ! the abstract interface excess_energy_t exactly); the explicit-size TARGET dummy here
    ! takes the address of the contiguous actual array.
    subroutine to_device(a, dst)
        real(dp), intent(in), target :: a(*)
        type(c_ptr), intent(in) :: dst
        call check(mdft_memcpy_h2d(ctx, dst, c_loc(a), 8_c_size_t * nvar), "upload")
    end subroutine

    subroutine from_device(src, a)
        type(c_ptr), intent(in) :: src
        real(dp), intent(inout), target :: a(*)
        call check(mdft_memcpy_d2h(ctx, c_loc(a), src, 8_c_size_t * nvar), "download")
    end subroutine

    ! The type B and type C corrections of a solute (atoms(1:6, natom): x, y, z, q, sigma, eps)
    ! in a solvent (sites(1:6, nsite)) for lattice electrostatics; a caller using the direct
    ! Coulomb sum (the reference's direct_sum = T) leaves them at 0.
    subroutine set_charged_solute_corrections(lengths, n0, atoms, sites, eps_r)
        real(dp), intent(in) :: lengths(3), n0, atoms(:, :), sites(:, :), eps_r
        real(dp) :: out(2)
        call check(mdft_charged_solute_corrections(lengths, n0, int(size(atoms, 2), c_int), atoms, &
                                                   int(size(sites, 2), c_int), sites, eps_r, out), &
                   "charged_solute_corrections")
        pbc_correction = out(1)
        pscheme_correction = out(2)
    end subroutine

    ! excess_energy_t-conformant plug-in (module_proc_ptrs.f90:3-21): df accumulated; it
    ! can be bound to a procedure pointer of that abstract interface, as init_vexc binds
    ! energy_cproj_mrso (src/module_init_simu.f90:60-73).
    subroutine excess_energy_gpu(ff, df, print_timers)
        real(dp), intent(out) :: ff
        real(dp), contiguous, intent(inout), optional :: df(:,:,:,:,:)
        logical, intent(in), optional :: print_timers
        if (present(df)) then
            call to_device(df, df_dev)
            call check(mdft_excess_energy(ctx, xi_dev, ff, df_dev, MDFT_DEVICE_PTRS), "excess_energy")
            call from_device(df_dev, df)
        else
            call check(mdft_excess_energy(ctx, xi_dev, ff, c_null_ptr, MDFT_DEVICE_PTRS), "excess_energy")
        end if
        if (present(print_timers)) then
            if (print_timers) print '(a,es16.8)', "excess energy (kJ/mol): ", ff
        end if
    end subroutine

end module mdft_gpu
