! ref_bridge.f90 -- bind(C) entry points onto the reference's own pure
! modules module_wigner_d and module_rotation, compiled from the reference
! sources by oracle/Makefile into oracle/_ref/.  TEST INFRASTRUCTURE ONLY:
! used by tests to check the oracle's restatement of these two functions.
module mdft_ref_bridge
    use iso_c_binding, only: c_int, c_double, c_double_complex
    implicit none
contains
    function ref_wigner_small_d(m, mu, mup, theta) bind(C, name="ref_wigner_small_d") result(d)
        use module_wigner_d, only: wigner_small_d
        integer(c_int), value :: m, mu, mup
        real(c_double), value :: theta
        real(c_double) :: d
        d = wigner_small_d(int(m), int(mu), int(mup), real(theta, 8))
    end function

    ! R(0:mmax,-mmax:mmax,-mmax:mmax) is returned in C order [m][mup][khi]
    subroutine ref_rotation_matrix(mmax, q, out) bind(C, name="ref_rotation_matrix")
        use module_rotation, only: rot => rotation_matrix_between_complex_spherical_harmonics_lu, rinit => init
        integer(c_int), value :: mmax
        real(c_double), intent(in) :: q(3)
        complex(c_double_complex), intent(out) :: out(*)
        complex(8), allocatable :: r(:,:,:)
        integer :: m, a, b, k, w
        logical, save :: ready = .false.
        if (.not. ready) then
            call rinit
            ready = .true.
        end if
        allocate(r(0:mmax,-mmax:mmax,-mmax:mmax))
        r = rot(int(mmax), q)
        w = 2*mmax+1
        k = 0
        do m = 0, mmax
            do a = -mmax, mmax
                do b = -mmax, mmax
                    k = k + 1
                    out(k) = r(m, a, b)
                end do
            end do
        end do
    end subroutine
end module mdft_ref_bridge
