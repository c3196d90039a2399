! mdft_lbfgs.f90 -- device-resident L-BFGS minimisation of the MDFT functional.
!
! Replaces the minimisation loop of src/energy_minimization.f90:232-290 and the
! reverse-communication L-BFGS-B of src/module_lbfgs_nocedal_mdft.f90 (used
! there unconstrained, with memory m = 1) for the device path: all length-n
! vectors (xi, the gradient, the search direction, the m correction pairs)
! stay in HBM and are combined with the library's FP64 vector operations; the
! scalars of the two-loop recursion and of the line search are FP64 on the
! host.  The line search enforces the strong Wolfe conditions with cubic
! interpolation inside a bracket (the role of dcsrch,
! src/module_lbfgs_nocedal_mdft.f90:1653).  Stopping rule: |dF|/|dF_0| < gtol,
! or a relative decrease of F below ftol, or maxeval evaluations.
module mdft_lbfgs
    use iso_c_binding
    use mdft_gpu, only: mdft_energy_and_gradient, mdft_device_alloc, mdft_device_free, mdft_vec_dot, &
                        mdft_vec_axpy_dot, mdft_vec_lincomb, mdft_get_info, MDFT_DEVICE_PTRS
    implicit none
    private
    integer, parameter :: dp = c_double
    integer, parameter :: mmem_max = 32
    public :: mdft_minimize

contains

    ! Minimise F over xi (device vector of length n), starting from the values in xi.
    ! On return xi holds the last accepted iterate and g its gradient.
    ! stats(1:4)  F, F_id, F_ext, F_exc at xi
    ! stats(5)    |g|/|g_0|
    ! stats(6)    evaluations used, stats(7) iterations, stats(8) status
    !             (0 converged on gtol, 1 converged on ftol, 2 maxeval, 3 line search failed,
    !              4 F within fband of ftarget when fband > 0)
    ! stats(9)    |g_0| (2-norm of the first gradient)
    ! trace(:, k) (when ntrace > 0) F, |g|/|g0| and evaluations used after each accepted iterate
    ! fnoise: relative noise level of F; the sufficient-decrease test is relaxed by
    ! fnoise*|F| so that steps are judged by the curvature condition once F stops
    ! resolving the decrease (the evaluation's single-precision chops put ~1e-8
    ! relative noise on F in FP64, more in the mixed-precision modes)
    integer(c_int) function mdft_minimize(ctx, n, xi, g, m, gtol, ftol, fnoise, ftarget, fband, maxeval, &
                                          stats, trace, ntrace) bind(C, name="mdft_minimize") result(rc)
        type(c_ptr), value :: ctx, xi, g
        integer(c_size_t), value :: n
        integer(c_int), value :: m, maxeval, ntrace
        real(c_double), value :: gtol, ftol, fnoise, ftarget, fband
        real(c_double), intent(out) :: stats(16)
        real(c_double), intent(out) :: trace(3, *)
        type(c_ptr) :: s(mmem_max), y(mmem_max), d, xn, gn, tmp, xi_user, g_user, xn_own, gn_own
        real(dp) :: rho(mmem_max), alpha(mmem_max)
        real(dp) :: parts(3), pn(3), f, fn, gnorm, g0, gd, gdn, t, dotv, beta, ys, yy, gam
        real(dp) :: f_prev
        integer :: k, i, j, nstored, newest, neval, iter, status, slot, mm
        integer(c_size_t) :: bytes
        logical :: ok

        rc = 0
        stats = 0
        mm = max(1, min(int(m), mmem_max))
        bytes = 8_c_size_t * n
        do i = 1, mm
            rc = mdft_device_alloc(ctx, bytes, s(i)); if (rc /= 0) return
            rc = mdft_device_alloc(ctx, bytes, y(i)); if (rc /= 0) return
        end do
        rc = mdft_device_alloc(ctx, bytes, d); if (rc /= 0) return
        rc = mdft_device_alloc(ctx, bytes, xn); if (rc /= 0) return
        rc = mdft_device_alloc(ctx, bytes, gn); if (rc /= 0) return
        xi_user = xi
        g_user = g
        xn_own = xn
        gn_own = gn

        neval = 0
        iter = 0
        status = 2
        rc = mdft_energy_and_gradient(ctx, xi, parts, g, MDFT_DEVICE_PTRS); if (rc /= 0) goto 900
        neval = 1
        f = sum(parts)
        rc = mdft_vec_dot(ctx, n, g, g, dotv); if (rc /= 0) goto 900
        g0 = sqrt(dotv)
        gnorm = g0
        nstored = 0
        newest = 0
        gam = 1.0_dp
        f_prev = huge(1.0_dp)
        if (g0 == 0.0_dp) then
            status = 0
            goto 800
        end if

        do while (neval < maxeval)
            if (gnorm / g0 < gtol) then
                status = 0
                exit
            end if
            ! ---- two-loop recursion: d = H g (search direction is -d)
            rc = mdft_vec_lincomb(ctx, n, 1.0_dp, g, 0.0_dp, c_null_ptr, d); if (rc /= 0) goto 900
            if (nstored > 0) then
                rc = mdft_vec_dot(ctx, n, s(newest), d, dotv); if (rc /= 0) goto 900
                do j = 0, nstored - 1
                    slot = modulo(newest - 1 - j, mm) + 1
                    alpha(slot) = rho(slot) * dotv
                    if (j < nstored - 1) then
                        i = modulo(slot - 2, mm) + 1
                        rc = mdft_vec_axpy_dot(ctx, n, -alpha(slot), y(slot), d, s(i), dotv)
                    else
                        rc = mdft_vec_axpy_dot(ctx, n, -alpha(slot), y(slot), d, c_null_ptr, dotv)
                    end if
                    if (rc /= 0) goto 900
                end do
                rc = mdft_vec_lincomb(ctx, n, gam, d, 0.0_dp, c_null_ptr, d); if (rc /= 0) goto 900
                slot = modulo(newest - nstored, mm) + 1
                rc = mdft_vec_dot(ctx, n, y(slot), d, dotv); if (rc /= 0) goto 900
                do j = nstored - 1, 0, -1
                    slot = modulo(newest - 1 - j, mm) + 1
                    beta = rho(slot) * dotv
                    if (j > 0) then
                        i = modulo(slot, mm) + 1
                        rc = mdft_vec_axpy_dot(ctx, n, alpha(slot) - beta, s(slot), d, y(i), dotv)
                    else
                        rc = mdft_vec_axpy_dot(ctx, n, alpha(slot) - beta, s(slot), d, c_null_ptr, dotv)
                    end if
                    if (rc /= 0) goto 900
                end do
                t = 1.0_dp
            else
                t = 1.0_dp / gnorm       ! first step: unit-length move along -g
            end if
            rc = mdft_vec_dot(ctx, n, g, d, gd); if (rc /= 0) goto 900
            gd = -gd                     ! directional derivative along -d
            if (gd >= 0.0_dp) then       ! not a descent direction: restart from steepest descent
                nstored = 0
                rc = mdft_vec_lincomb(ctx, n, 1.0_dp, g, 0.0_dp, c_null_ptr, d); if (rc /= 0) goto 900
                gd = -gnorm * gnorm
                t = 1.0_dp / gnorm
            end if
            call line_search(ok)
            if (rc /= 0) goto 900
            if (.not. ok) then
                if (nstored > 0) then    ! forget the memory once and retry along -g
                    nstored = 0
                    cycle
                end if
                status = 3
                exit
            end if
            ! ---- update the correction pairs: s = xn - x = -t d, y = gn - g
            newest = modulo(newest, mm) + 1
            rc = mdft_vec_lincomb(ctx, n, -t, d, 0.0_dp, c_null_ptr, s(newest)); if (rc /= 0) goto 900
            rc = mdft_vec_lincomb(ctx, n, 1.0_dp, gn, -1.0_dp, g, y(newest)); if (rc /= 0) goto 900
            rc = mdft_vec_dot(ctx, n, y(newest), s(newest), ys); if (rc /= 0) goto 900
            rc = mdft_vec_dot(ctx, n, y(newest), y(newest), yy); if (rc /= 0) goto 900
            if (ys > 1.0e-300_dp) then
                rho(newest) = 1.0_dp / ys
                gam = ys / yy
                nstored = min(nstored + 1, mm)
            else
                newest = modulo(newest - 2, mm) + 1
            end if
            tmp = xi; xi = xn; xn = tmp
            tmp = g; g = gn; gn = tmp
            f_prev = f
            f = fn
            parts = pn
            rc = mdft_vec_dot(ctx, n, g, g, dotv); if (rc /= 0) goto 900
            gnorm = sqrt(dotv)
            iter = iter + 1
            if (iter <= ntrace) then
                trace(1, iter) = f
                trace(2, iter) = gnorm / g0
                trace(3, iter) = neval
            end if
            if (fband > 0.0_dp .and. abs(f - ftarget) <= fband) then
                status = 4
                exit
            end if
            if (abs(f_prev - f) <= ftol * max(abs(f_prev), abs(f), 1.0_dp)) then
                status = 1
                exit
            end if
        end do
        if (gnorm / g0 < gtol .and. status /= 4) status = 0

800     continue
        stats(1) = f
        stats(2:4) = parts
        stats(5) = gnorm / max(g0, tiny(1.0_dp))
        stats(6) = neval
        stats(7) = iter
        stats(8) = status
        stats(9) = g0
900     continue
        ! the caller's xi/g buffers must hold the result
        if (rc == 0 .and. .not. c_associated(xi, xi_user)) rc = mdft_vec_lincomb(ctx, n, 1.0_dp, xi, 0.0_dp, c_null_ptr, xi_user)
        if (rc == 0 .and. .not. c_associated(g, g_user)) rc = mdft_vec_lincomb(ctx, n, 1.0_dp, g, 0.0_dp, c_null_ptr, g_user)
        call release()

    contains

        subroutine line_search(ok)
            logical, intent(out) :: ok
            real(dp), parameter :: c1 = 1.0e-4_dp, c2 = 0.9_dp
            real(dp) :: tlo, thi, flo, fhi, glo, ghi, tt, epsf
            integer :: it
            logical :: bracketed
            ok = .false.
            epsf = fnoise * max(abs(f), 1.0_dp)
            tlo = 0.0_dp
            flo = f
            glo = gd
            thi = 0.0_dp
            fhi = 0.0_dp
            ghi = 0.0_dp
            bracketed = .false.
            do it = 1, 30
                if (neval >= maxeval) return
                rc = mdft_vec_lincomb(ctx, n, 1.0_dp, xi, -t, d, xn); if (rc /= 0) return
                rc = mdft_energy_and_gradient(ctx, xn, pn, gn, MDFT_DEVICE_PTRS); if (rc /= 0) return
                neval = neval + 1
                fn = sum(pn)
                rc = mdft_vec_dot(ctx, n, gn, d, gdn); if (rc /= 0) return
                gdn = -gdn
                if (fn > f + c1 * t * gd + epsf .or. (bracketed .and. fn >= flo + epsf)) then
                    thi = t; fhi = fn; ghi = gdn
                    bracketed = .true.
                else
                    if (abs(gdn) <= -c2 * gd) then
                        ok = .true.
                        return
                    end if
                    if (gdn * (thi - tlo) >= 0.0_dp .and. bracketed) then
                        thi = tlo; fhi = flo; ghi = glo
                    end if
                    tlo = t; flo = fn; glo = gdn
                    if (.not. bracketed) then
                        if (gdn < 0.0_dp) then
                            t = 4.0_dp * t         ! still descending: extrapolate
                            cycle
                        end if
                        thi = tlo; fhi = flo; ghi = glo
                        bracketed = .true.
                    end if
                end if
                ! cubic interpolation of phi between tlo and thi, safeguarded
                tt = cubic_min(tlo, flo, glo, thi, fhi, ghi)
                if (.not. (tt > min(tlo, thi) + 0.1_dp * abs(thi - tlo) .and. &
                           tt < max(tlo, thi) - 0.1_dp * abs(thi - tlo))) tt = 0.5_dp * (tlo + thi)
                t = tt
                if (abs(thi - tlo) <= 1.0e-12_dp * max(1.0_dp, t)) exit
            end do
            ! accept the best Armijo point found, if any
            if (tlo > 0.0_dp .and. flo < f) then
                t = tlo
                rc = mdft_vec_lincomb(ctx, n, 1.0_dp, xi, -t, d, xn); if (rc /= 0) return
                rc = mdft_energy_and_gradient(ctx, xn, pn, gn, MDFT_DEVICE_PTRS); if (rc /= 0) return
                neval = neval + 1
                fn = sum(pn)
                ok = .true.
            end if
        end subroutine

        pure real(dp) function cubic_min(a, fa, ga, b, fb, gb)
            real(dp), intent(in) :: a, fa, ga, b, fb, gb
            real(dp) :: d1, d2, rad
            d1 = ga + gb - 3.0_dp * (fa - fb) / (a - b)
            rad = d1 * d1 - ga * gb
            if (rad < 0.0_dp .or. a == b) then
                cubic_min = 0.5_dp * (a + b)
                return
            end if
            d2 = sign(sqrt(rad), b - a)
            cubic_min = b - (b - a) * (gb + d2 - d1) / (gb - ga + 2.0_dp * d2)
        end function

        subroutine release()
            integer :: q
            integer(c_int) :: r2
            do q = 1, mm
                r2 = mdft_device_free(ctx, s(q))
                r2 = mdft_device_free(ctx, y(q))
            end do
            r2 = mdft_device_free(ctx, d)
            r2 = mdft_device_free(ctx, gn_own)
            r2 = mdft_device_free(ctx, xn_own)
        end subroutine

    end function mdft_minimize

end module mdft_lbfgs
