!> ref_driver.f90 -- C-callable entry points into the reference's own Fortran
!! modules `utils` (pysofft/utils.f90) and `make_wigner` (pysofft/wigner.f90),
!! compiled from the reference sources where they lie (see oracle/Makefile).
!!
!! TEST INFRASTRUCTURE ONLY: used by tests/ and tests/golden/make_golden.py to pin
!! the oracle; never linked into the product library.  The transform engine
!! (soft.f08) is not built: it needs FFTW3, which this image does not have.
!! All arrays are explicit-shape; matrices are returned in C order.
module ref_driver
  use, intrinsic :: iso_c_binding
  use precision
  use utils
  use make_wigner
  implicit none
contains
  function ref_total_num_coeffs(bw) bind(C, name="ref_total_num_coeffs") result(n)
    integer(c_int64_t), value :: bw
    integer(c_int64_t) :: n
    n = total_num_coeffs(bw)
  end function

  !> utils.f90:122-133, 1-based index of f^l_{m1,m2}
  function ref_coeff_location_mnl(m1, m2, l, bw) bind(C, name="ref_coeff_location_mnl") result(id)
    integer(c_int64_t), value :: m1, m2, l, bw
    integer(c_int64_t) :: id, a, b, c, d
    a = m1; b = m2; c = l; d = bw
    id = coeff_location_mnl(a, b, c, d)
  end function

  !> utils.f90:140-201, written as C-order [NC][3] = (l, m, n)
  subroutine ref_coeff_degrees(bw, out) bind(C, name="ref_coeff_degrees")
    integer(c_int64_t), value :: bw
    integer(c_int64_t), intent(out) :: out(3, (4*bw*bw*bw - bw)/3)
    integer(c_int64_t) :: lmn((4*bw*bw*bw - bw)/3, 3)
    lmn = get_coeff_degrees(bw)
    out = transpose(lmn)
  end subroutine

  !> utils.f90:389-404
  subroutine ref_legendre_weights(bw, w) bind(C, name="ref_legendre_weights")
    integer(c_int64_t), value :: bw
    real(c_double), intent(out) :: w(2*bw)
    w = legendre_quadrature_weights(bw)
  end subroutine

  !> wigner.f90:34-39
  subroutine ref_beta_samples(n, b) bind(C, name="ref_beta_samples")
    integer(c_int64_t), value :: n
    real(c_double), intent(out) :: b(n)
    b = create_beta_samples(n)
  end subroutine

  !> wigner.f90:94-100
  subroutine ref_alpha_gamma_samples(n, a) bind(C, name="ref_alpha_gamma_samples")
    integer(c_int64_t), value :: n
    real(c_double), intent(out) :: a(n)
    a = create_alpha_gamma_samples(n)
  end subroutine

  !> wigner.f90:164-187 with the trig samples of wigner.f90:51-58, C order [tri][2bw]
  subroutine ref_dlml_all(bw, out) bind(C, name="ref_dlml_all")
    integer(c_int64_t), value :: bw
    real(c_double), intent(out) :: out(2*bw, (bw*(bw + 1))/2)
    real(c_double), allocatable :: trig(:, :)
    integer(c_int64_t) :: b
    b = bw
    allocate(trig(2*bw, 3))
    trig = create_trig_samples(b)
    out = compute_all_dlml_l_contiguous(b, trig(:, 2), trig(:, 3), .true.)
  end subroutine

  !> wigner.f90:317-345 genwig_l2_trsp on the bw grid, C order [l-m2][2bw]
  subroutine ref_genwig(m1, m2, bw, out) bind(C, name="ref_genwig")
    integer(c_int64_t), value :: m1, m2, bw
    real(c_double), intent(out) :: out(2*bw, bw - m2)
    real(c_double), allocatable :: trig(:, :), dlml(:, :), w(:, :)
    integer(c_int64_t) :: b, a1, a2
    b = bw; a1 = m1; a2 = m2
    allocate(trig(2*bw, 3), dlml(2*bw, (bw*(bw + 1))/2), w(bw - m2, 2*bw))
    trig = create_trig_samples(b)
    dlml = compute_all_dlml_l_contiguous(b, trig(:, 2), trig(:, 3), .true.)
    w = genwig_l2_trsp(a1, a2, b, trig(:, 1), dlml(:, triangular_to_flat_index(a1, a2, b)), .true.)
    out = transpose(w)
  end subroutine

  !> wigner.f90:358-417, result C order [beta][2l+1][2l+1] = Fortran dl(beta, n, m)
  subroutine ref_wigner_dl_kostelec(l, nb, betas, normalized, out) bind(C, name="ref_wigner_dl_kostelec")
    integer(c_int64_t), value :: l, nb
    real(c_double), intent(in) :: betas(nb)
    integer(c_int), value :: normalized
    real(c_double), intent(out) :: out(2*l + 1, 2*l + 1, nb)
    real(c_double) :: dl(nb, 2*l + 1, 2*l + 1)
    integer(c_int64_t) :: i, j, ll
    ll = l
    dl = wigner_dl_kostelec(ll, betas, normalized /= 0)
    do i = 1, 2*l + 1
      do j = 1, 2*l + 1
        out(j, i, :) = dl(:, i, j)
      end do
    end do
  end subroutine

  !> wigner.f90:529-585 one Risbo step l-1 -> l for one beta sample
  subroutine ref_risbo_step(dprev, l, cb, sb, sqrts, nsq, dout) bind(C, name="ref_risbo_step")
    integer(c_int64_t), value :: l, nsq
    real(c_double), intent(in) :: dprev(max(1_c_int64_t, (2*l - 1)*(2*l - 1)))
    real(c_double), value :: cb, sb
    real(c_double), intent(in) :: sqrts(nsq)
    real(c_double), intent(out) :: dout((2*l + 1)*(2*l + 1))
    real(c_double) :: dp1(1, max(1_c_int64_t, (2*l - 1)*(2*l - 1))), r(1, (2*l + 1)*(2*l + 1))
    real(c_double) :: c1(1), s1(1)
    integer(c_int64_t) :: ll
    ll = l
    dp1(1, :) = dprev
    c1(1) = cb; s1(1) = sb
    if (l == 0) then
      r = wigner_mn_recurrence_risbo(dp1(:, 1:0), ll, c1, s1, sqrts)
    else
      r = wigner_mn_recurrence_risbo(dp1, ll, c1, s1, sqrts)
    end if
    dout = r(1, :)
  end subroutine

  !> wigner.f90:844-929, normalized D~^l_{mn}(alpha,beta,gamma)/(2 pi) in mnl order
  subroutine ref_full_big_d(bw, alpha, beta, gamma, out) bind(C, name="ref_full_big_d")
    integer(c_int64_t), value :: bw
    real(c_double), value :: alpha, beta, gamma
    complex(c_double_complex), intent(out) :: out((4*bw*bw*bw - bw)/3)
    integer(c_int64_t) :: b
    b = bw
    out = full_big_d_risbo(b, alpha, beta, gamma, .true.)
  end subroutine

  !> wigner.f90:844-934 full_big_d_risbo with the normalized flag
  subroutine ref_full_big_d_norm(bw, alpha, beta, gamma, normalized, out) bind(C, name="ref_full_big_d_norm")
    integer(c_int64_t), value :: bw
    real(c_double), value :: alpha, beta, gamma
    integer(c_int), value :: normalized
    complex(c_double_complex), intent(out) :: out((4*bw*bw*bw - bw)/3)
    integer(c_int64_t) :: b
    b = bw
    out = full_big_d_risbo(b, alpha, beta, gamma, normalized /= 0)
  end subroutine

  !> wigner.f90:626-748 one step of the symmetry-reduced Risbo recurrence, degree l-1 -> l,
  !! at nb betas (a set closed under beta -> pi - beta, reversed order); arrays C order
  !! [tri][nb] (= Fortran (nb, tri))
  subroutine ref_risbo_reduced_step(dl1, l, nb, cb, sb, sqrts, nsq, normalized, out) &
      bind(C, name="ref_risbo_reduced_step")
    integer(c_int64_t), value :: l, nb, nsq
    real(c_double), intent(in) :: dl1(nb, max(1_c_int64_t, (l*(l + 1))/2))
    real(c_double), intent(in) :: cb(nb), sb(nb), sqrts(nsq)
    integer(c_int), value :: normalized
    real(c_double), intent(out) :: out(nb, ((l + 1)*(l + 2))/2)
    integer(c_int64_t) :: ll
    ll = l
    if (l == 0) then
      out = wigner_recurrence_risbo_reduced(dl1(:, 1:0), ll, cb, sb, sqrts, normalized /= 0)
    else
      out = wigner_recurrence_risbo_reduced(dl1, ll, cb, sb, sqrts, normalized /= 0)
    end if
  end subroutine

  !> wigner.f90:784-822 genwig_all_risbo: every d~^l_{m,n}(beta_k), 0 <= m <= n <= l < bw, in
  !! mnl order, C order [mnl][2bw]
  subroutine ref_genwig_all_risbo(bw, normalized, out) bind(C, name="ref_genwig_all_risbo")
    integer(c_int64_t), value :: bw
    integer(c_int), value :: normalized
    real(c_double), intent(out) :: out(2*bw, (bw*(2 + 3*bw + bw*bw))/6)
    integer(c_int64_t) :: b
    b = bw
    out = genwig_all_risbo(b, normalized /= 0)
  end subroutine

  !> wigner.f90:430-491 genwig_all_kostelec, C order [mnl][2bw]
  subroutine ref_genwig_all_kostelec(bw, out) bind(C, name="ref_genwig_all_kostelec")
    integer(c_int64_t), value :: bw
    real(c_double), intent(out) :: out(2*bw, (bw*(2 + 3*bw + bw*bw))/6)
    integer(c_int64_t) :: b
    b = bw
    call genwig_all_kostelec_preallocated(b, out, .true.)
  end subroutine

  !> utils.f90:471-488
  subroutine ref_enforce_real_sym_mnl(coeff, bw) bind(C, name="ref_enforce_real_sym_mnl")
    integer(c_int64_t), value :: bw
    complex(c_double_complex), intent(inout) :: coeff((4*bw*bw*bw - bw)/3)
    call enforce_real_sym_mnl(coeff, bw)
  end subroutine

  !> utils.f90:261-292, 1-based index of f^l_{n1,n2} in lmn order
  function ref_coeff_location_lmn(l, n1, n2) bind(C, name="ref_coeff_location_lmn") result(id)
    integer(c_int64_t), value :: l, n1, n2
    integer(c_int64_t) :: id, a, b, c
    a = l; b = n1; c = n2
    id = coeff_location_lmn(a, b, c)
  end function

  !> utils.f90:298-355, written as C-order [NC][3] = (l, m, n)
  subroutine ref_coeff_degrees_risbo(bw, out) bind(C, name="ref_coeff_degrees_risbo")
    integer(c_int64_t), value :: bw
    integer(c_int64_t), intent(out) :: out(3, (4*bw*bw*bw - bw)/3)
    integer(c_int64_t) :: lmn((4*bw*bw*bw - bw)/3, 3)
    lmn = get_coeff_degrees_risbo(bw)
    out = transpose(lmn)
  end subroutine

  !> utils.f90:510-550
  subroutine ref_enforce_real_sym_lmn(coeff, bw) bind(C, name="ref_enforce_real_sym_lmn")
    integer(c_int64_t), value :: bw
    complex(c_double_complex), intent(inout) :: coeff((4*bw*bw*bw - bw)/3)
    call enforce_real_sym_lmn(coeff, bw)
  end subroutine

  !> utils.f90:735-741, 758-764, 781-787 (1-based)
  function ref_lmc(l, m) bind(C, name="ref_lmc") result(i)
    integer(c_int64_t), value :: l, m
    integer(c_int64_t) :: i, a, b
    a = l; b = m
    i = LMc(a, b)
  end function
  function ref_mlr(m, l, bw) bind(C, name="ref_mlr") result(i)
    integer(c_int64_t), value :: m, l, bw
    integer(c_int64_t) :: i, a, b, c
    a = m; b = l; c = bw
    i = MLr(a, b, c)
  end function
  function ref_mlc(m, l, bw) bind(C, name="ref_mlc") result(i)
    integer(c_int64_t), value :: m, l, bw
    integer(c_int64_t) :: i, a, b, c
    a = m; b = l; c = bw
    i = MLc(a, b, c)
  end function
end module ref_driver
