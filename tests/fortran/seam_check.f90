! seam_check.f90 -- TEST DRIVER of the Fortran drop-in seam (module mdft_gpu).
!
! Calls the reference-signature entry points the way the reference's own code
! calls its functional:
!   energy_and_gradient(f, df) with df(no,nx,ny,nz,1)
!       (src/module_energy_and_gradient.f90:36,63-64), and energy_and_gradient(f)
!       without df (src/module_pressure_correction.f90:124);
!   excess_energy_gpu bound to a procedure pointer of the reference's abstract
!       interface excess_energy_t (src/module_proc_ptrs.f90:3-21), as
!       init_vexc binds energy_cproj_mrso (src/module_init_simu.f90:60-73), and
!       called through it with df accumulated.
! Input (argument 1, stream file): int32 nx ny nz mmax mrso precision np_file nq_file,
! FP64 lx ly lz kT n0 dq, int32 m n mu nu khi (np_file each), FP64 c (2 np_file nq_file,
! interleaved), vext and xi (no nx ny nz each).  Output (argument 2, stream file): FP64
! f, f (F only), f_id, f_ext, f_exc, ff (through the pointer), df, df2 (1 + excess df).
program seam_check
    use iso_c_binding
    use mdft_gpu
    implicit none
    abstract interface
        subroutine excess_energy_t(ff, df, print_timers)
            import :: c_double
            real(c_double), intent(out) :: ff
            real(c_double), contiguous, intent(inout), optional :: df(:,:,:,:,:)
            logical, intent(in), optional :: print_timers
        end subroutine
    end interface
    procedure(excess_energy_t), pointer :: excess_energy => null()
    type(mdft_desc) :: d
    integer(c_int) :: hdr(8)
    real(c_double) :: rh(6)
    integer(c_int), allocatable :: tm(:), tn(:), tmu(:), tnu(:), tkhi(:)
    real(c_double), allocatable, target :: c(:), vext(:), xi(:)
    real(c_double), allocatable :: df(:,:,:,:,:), df2(:,:,:,:,:)
    real(c_double) :: f, f2, ff
    integer :: u, no, npf, nqf
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=trim(fin), access="stream", form="unformatted", status="old", action="read")
    read (u) hdr, rh
    npf = hdr(7)
    nqf = hdr(8)
    allocate (tm(npf), tn(npf), tmu(npf), tnu(npf), tkhi(npf), c(2 * npf * nqf))
    read (u) tm, tn, tmu, tnu, tkhi, c
    d = mdft_desc(hdr(1), hdr(2), hdr(3), rh(1), rh(2), rh(3), hdr(4), hdr(5), rh(4), rh(5), hdr(6), 0)
    call gpu_init(d)
    no = int(nvar / (int(hdr(1), c_size_t) * hdr(2) * hdr(3)))
    allocate (vext(nvar), xi(nvar))
    read (u) vext, xi
    close (u)
    call check(mdft_set_ck(ctx, npf, tm, tn, tmu, tnu, tkhi, nqf, rh(6), c), "set_ck")
    call check(mdft_set_vext(ctx, c_loc(vext), 0_c_int), "set_vext")
    call check(mdft_memcpy_h2d(ctx, xi_dev, c_loc(xi), 8_c_size_t * nvar), "upload xi")

    allocate (df(no, hdr(1), hdr(2), hdr(3), 1))
    call energy_and_gradient(f, df)
    call energy_and_gradient(f2)
    excess_energy => excess_energy_gpu
    allocate (df2(no, hdr(1), hdr(2), hdr(3), 1), source=1.0_c_double)
    call excess_energy(ff, df2, print_timers=.false.)

    open (newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace", action="write")
    write (u) f, f2, f_id, f_ext, f_exc, ff, df, df2
    close (u)
    call gpu_finalize()
end program seam_check
