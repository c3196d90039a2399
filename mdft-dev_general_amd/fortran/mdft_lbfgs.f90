! mdft_lbfgs.f90 -- device-resident L-BFGS minimisation of the MDFT functional.
!
! Replaces the minimisation loop of src/energy_minimization.f90:232-290 and the
! reverse-communication L-BFGS-B of src/module_lbfgs_nocedal_mdft.f90 (used
! there unconstrained, with memory m = 1) for the device path: all length-n
! vectors stay in HBM.  The recursion is the vector-free form of L-BFGS (Chen,
! Wang, Zhou & Gao, NIPS 2014): the correction pairs are kept implicitly as the
! search directions d_i (s_i = -t_i d_i) and the gradients g_i (y_i = g_{i+1} - g_i),
! their FP64 Gram matrix is kept on the host, and the two-loop recursion runs on
! coefficient vectors of length <= 2m+3.  Per iteration the device does one
! pass forming the new direction from <= 2m+1 vectors together with the first
! trial point, and one multi-dot of the new gradient with the <= 2m+3 live
! vectors that stay live (plus one step update per further trial) -- 4m + 5
! passes over n (9 at m = 1, 25 at m = 5) instead of the ~66 of the two-loop
! recursion on vectors.  The line search enforces the strong Wolfe conditions with cubic
! interpolation inside a bracket (the role of dcsrch,
! src/module_lbfgs_nocedal_mdft.f90:1653).  Stopping rule: |dF|/|dF_0| < gtol,
! or a decrease of F below ftol in one iteration (relative, or absolute as the
! reference's (fold - f) <= tol, src/module_lbfgs_nocedal_mdft.f90:425-427), or
! maxeval evaluations.
module mdft_lbfgs
    use iso_c_binding
    use mdft_gpu, only: mdft_energy_and_gradient, mdft_energy_and_gradient_dots, mdft_device_alloc, mdft_device_free, mdft_vec_lincomb, &
                        mdft_vec_multi_dot, mdft_vec_combine_step, mdft_vec_multi_dot_f32, &
                        mdft_vec_combine_step_f32, mdft_vec_axpy_f32, mdft_vec_widen_f32, &
                        mdft_scale_variables, mdft_vec_narrow_f64, mdft_vec_axpy_f32x, &
                        mdft_vec_combine_step_f32x, MDFT_DEVICE_PTRS, MDFT_DF_FP32, MDFT_SCALED_VARS, MDFT_XI_FP32
    implicit none
    private
    integer, parameter :: dp = c_double
    integer, parameter :: mmem_max = 6           ! 2 m + 3 live vectors per multi-dot <= 16
    integer, parameter :: nslot = 2 * mmem_max + 3
    public :: mdft_minimize, mdft_minimize_ex

contains

    ! mdft_minimize_ex with FP64 directions and gradients (the reference's precision)
    integer(c_int) function mdft_minimize(ctx, n, xi, g, m, gtol, ftol, fnoise, ftarget, fband, maxeval, &
                                          stats, trace, ntrace) bind(C, name="mdft_minimize") result(rc)
        type(c_ptr), value :: ctx, xi, g
        integer(c_size_t), value :: n
        integer(c_int), value :: m, maxeval, ntrace
        real(c_double), value :: gtol, ftol, fnoise, ftarget, fband
        real(c_double), intent(out) :: stats(16)
        real(c_double), intent(out) :: trace(3, *)
        rc = mdft_minimize_ex(ctx, n, xi, g, m, gtol, ftol, fnoise, ftarget, fband, maxeval, stats, trace, ntrace, 0, &
                              0.0_dp)
    end function

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
    ! opts bit 0 (1): the stored directions and gradients are FP32 (half the bytes of every
    ! vector pass; FP64 dot products, iterate and Gram matrix) -- for the mixed-precision
    ! evaluations, whose gradients carry FP32-level errors anyway.
    ! opts bit 1 (2): minimise over u = s xi, s(io) = sqrt(w(io) / mean w) (MDFT_SCALED_VARS):
    ! every term's curvature in xi carries the quadrature weight w(io), which spans a factor
    ! of about 4 over the theta nodes at mmax = 5; in u it does not.  xi is converted in
    ! place on entry and back on return.
    ! opts bit 3 (8), with bit 0: the iterate itself (x and the trial points) is stored in FP32
    ! too (MDFT_XI_FP32 evaluations): 8 fewer bytes per element in every iteration and
    ! evaluation; xi (FP64) receives the final iterate widened.
    ! opts bit 2 (4): ftol is absolute, the reference's stop (fold - f) <= tol
    ! (src/module_lbfgs_nocedal_mdft.f90:425-427, tol = factr * epsmch = 1e-3 kJ/mol by
    ! default, :10,188); without it the stop is |f_prev - f| <= ftol max(|f_prev|, |f|, 1).
    ! step0 > 0: the first trial step along -g of an iteration without curvature pairs (the
    ! first, or after a restart), in place of the unit-length move t = 1/|g|; a caller that
    ! knows the functional's curvature c along the variables passes 1/c, the Newton step.
    ! xi and g (FP64) receive the last iterate and its gradient dF/dxi in every mode.
    integer(c_int) function mdft_minimize_ex(ctx, n, xi, g, m, gtol, ftol, fnoise, ftarget, fband, maxeval, &
                                             stats, trace, ntrace, opts, step0) bind(C, name="mdft_minimize_ex") &
                                             result(rc)
        type(c_ptr), value :: ctx, xi, g
        integer(c_size_t), value :: n
        integer(c_int), value :: m, maxeval, ntrace, opts
        real(c_double), value :: step0
        real(c_double), value :: gtol, ftol, fnoise, ftarget, fband
        real(c_double), intent(out) :: stats(16)
        real(c_double), intent(out) :: trace(3, *)
        ! vector slots: 1..mm+1 hold directions, mm+2..2mm+3 gradients (slot mm+2 starts as the caller's g)
        type(c_ptr) :: vec(nslot), x, xn, tmp, xn_own, x32_own, pv(16)
        logical :: owned(nslot)
        real(dp) :: gram(nslot, nslot), tstep(nslot)
        integer :: ds(mmem_max + 1), gs(mmem_max + 2)   ! window: direction slots of pairs 1..k, gradient slots 1..k+1
        real(dp) :: q(nslot), r(nslot), sv(nslot), yv(nslot), alpha(mmem_max), rho(mmem_max)
        real(dp) :: coef(16), dots(16)
        integer :: lst(16)
        real(dp) :: parts(3), pn(3), f, fn, gnorm, g0, gd, gdn, t, beta, sy, yy, gam, f_prev
        real(dp) :: t_ready   ! step for which xn already holds x - t d (negative: none)
        integer :: i, j, k, mm, nb, nl, fd, fg, gcur, neval, iter, status
        integer(c_size_t) :: bytes
        integer(c_int) :: eflags, r2
        logical :: ok, f32, sc, scaled_in, fabs_stop, x32

        rc = 0
        stats = 0
        mm = max(1, min(int(m), mmem_max))
        gcur = mm + 2
        f32 = iand(opts, 1_c_int) /= 0
        sc = iand(opts, 2_c_int) /= 0
        fabs_stop = iand(opts, 4_c_int) /= 0
        x32 = f32 .and. iand(opts, 8_c_int) /= 0
        scaled_in = .false.
        bytes = merge(4_c_size_t, 8_c_size_t, f32) * n
        eflags = MDFT_DEVICE_PTRS
        if (f32) eflags = eflags + MDFT_DF_FP32
        if (sc) eflags = eflags + MDFT_SCALED_VARS
        if (x32) eflags = eflags + MDFT_XI_FP32
        owned = .false.
        xn_own = c_null_ptr
        x32_own = c_null_ptr
        t_ready = -1.0_dp
        do i = 1, 2 * mm + 3
            vec(i) = c_null_ptr
        end do
        do i = 1, 2 * mm + 3
            if (i == mm + 2 .and. .not. f32) then
                vec(i) = g
            else
                rc = mdft_device_alloc(ctx, bytes, vec(i)); if (rc /= 0) goto 900
                owned(i) = .true.
            end if
        end do
        rc = mdft_device_alloc(ctx, merge(4_c_size_t, 8_c_size_t, x32) * n, xn_own); if (rc /= 0) goto 900
        if (sc) then
            rc = mdft_scale_variables(ctx, n, 1_c_int, 0_c_int, xi, xi); if (rc /= 0) goto 900
            scaled_in = .true.
        end if
        x = xi
        xn = xn_own
        if (x32) then                            ! the iterate in FP32: x32_own = (float) xi
            rc = mdft_device_alloc(ctx, 4_c_size_t * n, x32_own); if (rc /= 0) goto 900
            rc = mdft_vec_narrow_f64(ctx, n, xi, x32_own); if (rc /= 0) goto 900
            x = x32_own
        end if
        gram = 0.0_dp
        tstep = 0.0_dp

        neval = 0
        iter = 0
        status = 2
        gcur = mm + 2
        pv(1) = vec(gcur)
        if (f32) then
            rc = mdft_energy_and_gradient_dots(ctx, x, parts, vec(gcur), eflags, 1_c_int, pv, dots)
        else
            rc = mdft_energy_and_gradient(ctx, x, parts, vec(gcur), eflags)
            if (rc == 0) rc = multi_dot(1)
        end if
        if (rc /= 0) goto 900
        neval = 1
        f = sum(parts)
        gram(gcur, gcur) = dots(1)
        g0 = sqrt(dots(1))
        gnorm = g0
        k = 0
        gs(1) = gcur
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
            ! ---- two-loop recursion on coefficients: r = H g (search direction is -r)
            q = 0.0_dp
            q(gcur) = 1.0_dp
            do i = k, 1, -1
                call pair(i, sv, yv)
                rho(i) = 1.0_dp / gdot(sv, yv)
                alpha(i) = rho(i) * gdot(sv, q)
                q = q - alpha(i) * yv
            end do
            gam = 1.0_dp
            if (k > 0) then
                call pair(k, sv, yv)
                gam = gdot(sv, yv) / gdot(yv, yv)
            end if
            r = gam * q
            do i = 1, k
                call pair(i, sv, yv)
                beta = rho(i) * gdot(yv, r)
                r = r + (alpha(i) - beta) * sv
            end do
            t = 1.0_dp
            if (k == 0) t = first_step()         ! first step along -g
            gd = -gdot(unit(gcur), r)            ! directional derivative along -d
            if (gd >= 0.0_dp) then               ! not a descent direction: restart from steepest descent
                k = 0
                gs(1) = gcur
                r = unit(gcur)
                gd = -gnorm * gnorm
                t = first_step()
            end if
            ! ---- the new direction d = sum_j r_j b_j into a free direction slot
            fd = free_direction()
            nb = 0
            do j = 1, 2 * mm + 3
                if (r(j) /= 0.0_dp) then
                    nb = nb + 1
                    pv(nb) = vec(j)
                    coef(nb) = r(j)
                end if
            end do
            ! the direction and the line search's first trial point in one pass
            if (x32) then
                rc = mdft_vec_combine_step_f32x(ctx, n, nb, coef, pv, vec(fd), -t, x, xn)
            else if (f32) then
                rc = mdft_vec_combine_step_f32(ctx, n, nb, coef, pv, vec(fd), -t, x, xn)
            else
                rc = mdft_vec_combine_step(ctx, n, nb, coef, pv, vec(fd), -t, x, xn)
            end if
            if (rc /= 0) goto 900
            t_ready = t
            do j = 1, 2 * mm + 3
                if (j == fd) cycle
                gram(fd, j) = dot_product(r, gram(:, j))
                gram(j, fd) = gram(fd, j)
            end do
            gram(fd, fd) = gdot(r, r)
            fg = free_gradient()
            call line_search(ok)
            if (rc /= 0) goto 900
            if (.not. ok) then
                if (k > 0) then                  ! forget the memory once and retry along -g
                    k = 0
                    gs(1) = gcur
                    cycle
                end if
                status = 3
                exit
            end if
            ! ---- accept: pair (s = -t d, y = g_new - g)
            sy = -t * (gram(fd, fg) - gram(fd, gcur))
            if (sy > 1.0e-300_dp) then
                tstep(fd) = t
                k = k + 1
                ds(k) = fd
                gs(k + 1) = fg
                if (k > mm) then                 ! drop the oldest pair
                    ds(1:k - 1) = ds(2:k)
                    gs(1:k) = gs(2:k + 1)
                    k = k - 1
                end if
            else
                k = 0
                gs(1) = fg
            end if
            gcur = fg
            tmp = x; x = xn; xn = tmp
            f_prev = f
            f = fn
            parts = pn
            gnorm = sqrt(gram(gcur, gcur))
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
            if (fabs_stop) then
                if (f_prev - f <= ftol) then
                    status = 1
                    exit
                end if
            else if (abs(f_prev - f) <= ftol * max(abs(f_prev), abs(f), 1.0_dp)) then
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
        if (x32 .and. scaled_in .and. rc == 0) then
            ! xi = u / s straight from the FP32 iterate: one pass instead of a widening and
            ! an unscaling of xi
            rc = mdft_scale_variables(ctx, n, -1_c_int, 1_c_int, x, xi)
            if (rc == 0) scaled_in = .false.
        else if (x32) then
            if (rc == 0) rc = mdft_vec_widen_f32(ctx, n, x, xi)
        else if (rc == 0 .and. .not. c_associated(x, xi)) then
            rc = mdft_vec_lincomb(ctx, n, 1.0_dp, x, 0.0_dp, c_null_ptr, xi)
        end if
        if (rc == 0 .and. sc .and. c_associated(vec(gcur))) then
            ! dF/dxi = s dF/du
            rc = mdft_scale_variables(ctx, n, 1_c_int, merge(1_c_int, 0_c_int, f32), vec(gcur), g)
        else if (rc == 0 .and. f32 .and. c_associated(vec(gcur))) then
            rc = mdft_vec_widen_f32(ctx, n, vec(gcur), g)
        else if (rc == 0 .and. .not. c_associated(vec(gcur), g)) then
            rc = mdft_vec_lincomb(ctx, n, 1.0_dp, vec(gcur), 0.0_dp, c_null_ptr, g)
        end if
        if (scaled_in) then                      ! xi = u / s, also after an error
            r2 = mdft_scale_variables(ctx, n, -1_c_int, 0_c_int, xi, xi)
            if (rc == 0) rc = r2
        end if
        call release()

    contains

        ! the step of an iteration without curvature information: step0, or a unit-length move
        real(dp) function first_step()
            first_step = 1.0_dp / gnorm
            if (step0 > 0.0_dp) first_step = step0
        end function

        ! dots(1:nl_) = pv(j) . pv(nl_): the last vector of pv is the new gradient itself
        integer(c_int) function multi_dot(nl_) result(r)
            integer, intent(in) :: nl_
            if (f32) then
                r = mdft_vec_multi_dot_f32(ctx, n, int(nl_, c_int), pv, pv(nl_), dots)
            else
                r = mdft_vec_multi_dot(ctx, n, int(nl_, c_int), pv, pv(nl_), dots)
            end if
        end function

        ! a^T G b over the slot coefficients
        real(dp) function gdot(a, b)
            real(dp), intent(in) :: a(nslot), b(nslot)
            gdot = dot_product(a, matmul(gram, b))
        end function

        function unit(j) result(e)
            integer, intent(in) :: j
            real(dp) :: e(nslot)
            e = 0.0_dp
            e(j) = 1.0_dp
        end function

        ! coefficients of s_i = -t_i d_i and y_i = g_{i+1} - g_i
        subroutine pair(i, a, b)
            integer, intent(in) :: i
            real(dp), intent(out) :: a(nslot), b(nslot)
            a = 0.0_dp
            b = 0.0_dp
            a(ds(i)) = -tstep(ds(i))
            b(gs(i + 1)) = b(gs(i + 1)) + 1.0_dp
            b(gs(i)) = b(gs(i)) - 1.0_dp
        end subroutine

        integer function free_direction()
            integer :: a
            do a = 1, mm + 1
                if (.not. any(ds(1:k) == a)) then
                    free_direction = a
                    return
                end if
            end do
            free_direction = mm + 1
        end function

        integer function free_gradient()
            integer :: a
            do a = mm + 2, 2 * mm + 3
                if (.not. any(gs(1:k + 1) == a)) then
                    free_gradient = a
                    return
                end if
            end do
            free_gradient = 2 * mm + 3
        end function

        ! evaluate at xn = x - t d into slot fg and its Gram row (one multi-dot with the live vectors)
        subroutine trial()
            integer :: a, a0
            if (t /= t_ready) then
                if (x32) then
                    rc = mdft_vec_axpy_f32x(ctx, n, -t, vec(fd), x, xn)
                else if (f32) then
                    rc = mdft_vec_axpy_f32(ctx, n, -t, vec(fd), x, xn)
                else
                    rc = mdft_vec_lincomb(ctx, n, 1.0_dp, x, -t, vec(fd), xn)
                end if
                if (rc /= 0) return
            end if
            t_ready = -1.0_dp
            ! with the memory full, the oldest pair (ds(1), gs(1)) is dropped after this
            ! iteration whatever the line search gives: its products with the new
            ! gradient are never used, so they are not computed
            a0 = 1
            if (k == mm) a0 = 2
            nl = 0
            do a = a0, k
                nl = nl + 1
                lst(nl) = ds(a)
            end do
            nl = nl + 1
            lst(nl) = fd
            do a = a0, k + 1
                nl = nl + 1
                lst(nl) = gs(a)
            end do
            nl = nl + 1
            lst(nl) = fg
            do a = 1, nl
                pv(a) = vec(lst(a))
            end do
            if (nl <= 3 .and. f32) then
                ! up to 3 products (m = 1) are formed by the evaluation as it writes the FP32
                ! gradient (with FP64 vectors the separate pass is the faster one)
                rc = mdft_energy_and_gradient_dots(ctx, xn, pn, vec(fg), eflags, int(nl, c_int), pv, dots)
                if (rc /= 0) return
            else
                rc = mdft_energy_and_gradient(ctx, xn, pn, vec(fg), eflags); if (rc /= 0) return
                rc = multi_dot(nl); if (rc /= 0) return
            end if
            neval = neval + 1
            fn = sum(pn)
            do a = 1, nl
                gram(fg, lst(a)) = dots(a)
                gram(lst(a), fg) = dots(a)
            end do
            gdn = -gram(fg, fd)
        end subroutine

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
                call trial(); if (rc /= 0) return
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
                call trial(); if (rc /= 0) return
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
            integer :: a
            integer(c_int) :: r2
            do a = 1, 2 * mm + 3
                if (owned(a)) r2 = mdft_device_free(ctx, vec(a))
            end do
            if (c_associated(xn_own)) r2 = mdft_device_free(ctx, xn_own)
            if (c_associated(x32_own)) r2 = mdft_device_free(ctx, x32_own)
        end subroutine

    end function mdft_minimize_ex

end module mdft_lbfgs
