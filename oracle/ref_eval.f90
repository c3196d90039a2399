! ref_eval.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A driver of our own around the reference's evaluation of the functional:
! energy_ideal_and_external (src/module_energy_ideal_and_external.f90:14-74)
! and energy_cproj_mrso (src/module_energy_cproj_mrso.f90:61-782), compiled
! from the reference's sources together with FFTW 3.3.4 built from the
! reference's vendored source tarball (oracle/Makefile, target `ref`).  It is
! the reference half of the golden vectors under tests/golden/ and the CPU
! baseline's "reference" timing; nothing in the product calls it.
!
! It does what energy_and_gradient (src/module_energy_and_gradient.f90:36-240)
! does for a neutral solute and the cproj functional: dF zeroed, the ideal +
! external term, then the excess term, each accumulating into dF.
!
! Run in a directory holding
!   dft.in                                   boxlen, boxnod, mmax, temperature, solvent spce
!   data/dcf/spce/spce-ck_nonzero_nmax<mmax>_ml   the c(k) file read_c_luc opens
!   output/                                  (read_c_luc writes a memory report there)
!   xi.bin, vext.bin                         raw FP64 in the reference layout (no, nx, ny, nz)
! Arguments: [nrep] [write_df] [inputs]
!   nrep      evaluations to run (the first one plans the FFTs and reads c(k));
!             the wall time of the later ones is reported per evaluation
!   write_df  1 (default): write df.bin (raw FP64, same layout); 0: F only
!   inputs    "files" (default): read xi.bin and vext.bin; "synthetic": fill
!             xi = 1 + 0.1 sin(.) and vext = 1 + cos(.) of the indices instead
!             (timing runs at full size without writing the fields to disk)
! Writes ref_out.txt: kT n0 rho0 dv, then F_id F_ext F_exc, then seconds per
! evaluation (first, later mean), then the OpenMP thread count.
program ref_eval
    use precision_kinds, only: dp
    use omp_lib, only: omp_get_max_threads
    use module_grid, only: grid
    use module_thermo, only: init_thermo, thermo
    use module_solvent, only: solvent, read_solvent
    use module_energy_ideal_and_external, only: energy_ideal_and_external
    use module_energy_cproj_mrso, only: energy_cproj_mrso
    implicit none
    real(dp), allocatable :: df(:,:,:,:,:)
    real(dp) :: fid, fext, fexc, t_first, t_rest
    integer :: no, nx, ny, nz, nrep, write_df, irep, u
    integer(8) :: c0, c1, rate
    character(len=32) :: arg, inputs
    integer :: io, ix, iy, iz

    nrep = 1
    write_df = 1
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read (arg, *) nrep
    end if
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read (arg, *) write_df
    end if
    inputs = "files"
    if (command_argument_count() >= 3) call get_command_argument(3, inputs)
    nrep = max(nrep, 1)

    call grid%init
    call init_thermo
    call read_solvent
    no = grid%no
    nx = grid%nx
    ny = grid%ny
    nz = grid%nz

    allocate (solvent(1)%xi(no, nx, ny, nz))
    allocate (solvent(1)%vext(no, nx, ny, nz))
    if (trim(inputs) == "synthetic") then
        !$omp parallel do private(ix, iy, io)
        do iz = 1, nz
            do iy = 1, ny
                do ix = 1, nx
                    do io = 1, no
                        solvent(1)%xi(io, ix, iy, iz) = 1._dp + 0.1_dp * sin(0.37_dp * io + 0.11_dp * ix &
                                                                             + 0.07_dp * iy + 0.05_dp * iz)
                        solvent(1)%vext(io, ix, iy, iz) = 1._dp + cos(0.29_dp * io + 0.13_dp * ix &
                                                                      + 0.03_dp * iy + 0.09_dp * iz)
                    end do
                end do
            end do
        end do
        !$omp end parallel do
    else
        open (newunit=u, file="xi.bin", access="stream", form="unformatted", status="old", action="read")
        read (u) solvent(1)%xi
        close (u)
        open (newunit=u, file="vext.bin", access="stream", form="unformatted", status="old", action="read")
        read (u) solvent(1)%vext
        close (u)
    end if
    allocate (df(no, nx, ny, nz, 1))

    t_rest = 0._dp
    t_first = 0._dp
    do irep = 1, nrep
        call system_clock(c0, rate)
        df = 0._dp
        call energy_ideal_and_external(fid, fext, df)
        call energy_cproj_mrso(fexc, df)
        call system_clock(c1)
        if (irep == 1) then
            t_first = real(c1 - c0, dp) / real(rate, dp)
        else
            t_rest = t_rest + real(c1 - c0, dp) / real(rate, dp)
        end if
    end do
    if (nrep > 1) t_rest = t_rest / real(nrep - 1, dp)

    open (newunit=u, file="ref_out.txt", status="replace", action="write")
    write (u, '(4ES25.17)') thermo%kbT, solvent(1)%n0, solvent(1)%rho0, grid%dv
    write (u, '(3ES25.17)') fid, fext, fexc
    write (u, '(2ES25.17)') t_first, t_rest
    write (u, '(I8)') omp_get_max_threads()
    close (u)
    if (write_df /= 0) then
        open (newunit=u, file="df.bin", access="stream", form="unformatted", status="replace", action="write")
        write (u) df
        close (u)
    end if
end program ref_eval
