!! Fortran host driving the gfx950 engine through the alpsg module, the way
!! the ALPS driver would: alps_var-layout arrays in, disp/secant_osc calls out.
!! The plasma is a proton-electron bi-Maxwellian tabulated with the formulas
!! of generate_distribution.f90 (distributions=1) on 24 x 48 points, kperp =
!! kpar = 0.3.  Prints nmax, det D at a few frequencies and one secant root;
!! then, at kperp = kpar = 0.05, the om_double_scan of a kpar scan to 0.2
!! (3 log steps) with a kperp scan to 0.2 (2 output steps of 2) from every
!! outer point, in place of om_double_scan (ALPS.f90:126); and two solves on
!! the root server with its statistics.
program alpsg_demo
  use iso_c_binding
  use alpsg
  implicit none
  integer, parameter :: nspec = 2, nperp = 24, npar = 48
  real(c_double), target :: pp(nspec, 0:nperp, 0:npar, 2), f0(nspec, 0:nperp, 0:npar)
  real(c_double), target :: ns(nspec), qs(nspec), ms(nspec), taus(nspec), alphs(nspec)
  real(c_double), target :: param_fit(nspec, 0:nperp, 5, 1), perp_correction(nspec, 1)
  integer(c_int), target :: fit_type(nspec, 1), ACmethod(nspec), n_fits(nspec), nmax(nspec)
  type(alpsg_plasma_desc) :: d
  type(alpsg_root_params) :: rp
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer :: is, iperp, ipar, k
  real(c_double) :: beta, vA, maxP, pi, norm, pperp_max, ppar_max, dpperp, dppar, pperp, ppar
  real(c_double) :: pcx
  complex(c_double_complex) :: om(4), det, root
  type(alpsg_scan_params) :: s1, s2
  integer, parameter :: nrow = (3 + 1) * (2 + 1)
  real(c_double) :: kpo(nrow), kzo(nrow), lat(2)
  complex(c_double_complex) :: droots(1), rout(nrow), g(2), sroots(2)
  integer(c_int) :: written(nrow), pid, iters(2), status(2), nout
  integer(c_long_long) :: tags(2), counts(10)
  real(c_double) :: times(4)
  type(c_ptr) :: srv

  pi = atan(1.d0) * 4.d0
  beta = 1.d0; vA = 1.d-4; maxP = 6.d0
  ms = [1.d0, 5.44662d-4]; qs = [1.d0, -1.d0]; ns = [1.d0, 1.d0]
  taus = [1.d0, 1.d0]; alphs = [1.d0, 1.d0]
  perp_correction(:, 1) = [1.d0 / (taus(1) * beta * ms(1) * alphs(1)), 1.836d3]
  do is = 1, nspec
     norm = pi**(-1.5d0) / ((ms(is) * beta * taus(is))**(3.d0 / 2.d0) * alphs(is))
     pperp_max = maxP * sqrt(ms(is) * taus(is) * alphs(is))
     ppar_max = maxP * sqrt(ms(is) * taus(is))
     dpperp = pperp_max / real(nperp)
     dppar = 2.d0 * ppar_max / real(npar)
     pcx = 1.d0 / (taus(is) * beta * ms(is) * alphs(is))
     do iperp = 0, nperp
        pperp = real(iperp) * dpperp
        do ipar = 0, npar
           ppar = real(ipar) * dppar - ppar_max
           f0(is, iperp, ipar) = exp(-((ppar**2.d0) / (beta * ms(is) * taus(is)) + &
                (pperp**2.d0) / (taus(is) * beta * ms(is) * alphs(is)))) * norm
           pp(is, iperp, ipar, 1) = pperp
           pp(is, iperp, ipar, 2) = ppar
        enddo
        ! exact Maxwell fit parameters (ALPS_analyt.f90:225-229)
        param_fit(is, iperp, 1, 1) = norm * exp((perp_correction(is, 1) - pcx) * pperp**2)
        param_fit(is, iperp, 2, 1) = 1.d0 / (beta * ms(is) * taus(is))
        param_fit(is, iperp, 3, 1) = 0.d0
        param_fit(is, iperp, 4:5, 1) = 0.d0
     enddo
  enddo
  fit_type = 1; ACmethod = 1; n_fits = 1

  d%nspec = nspec; d%nperp = nperp; d%npar = npar
  d%pp = c_loc(pp); d%f0 = c_loc(f0)
  d%ns = c_loc(ns); d%qs = c_loc(qs); d%ms = c_loc(ms)
  d%vA = vA; d%kperp_norm = 1; d%positions_principal = 3; d%n_resonance_interval = 100
  d%Tlim = 0.01d0; d%Bessel_zero = 1.d-45
  d%ACmethod = c_loc(ACmethod); d%n_fits = c_loc(n_fits); d%max_fits = 1
  d%fit_type = c_loc(fit_type); d%perp_correction = c_loc(perp_correction)
  d%param_fit = c_loc(param_fit); d%param_fit_nperp = nperp
  d%poly_max_order = 0

  rc = alpsg_create(d, h)
  if (rc /= ALPSG_OK) then
     write(*, '(a,i0,2a)') 'alpsg_create failed ', rc, ': ', trim(alpsg_error_text())
     stop 1
  endif
  rc = alpsg_set_k(h, 0.3d0, 0.3d0, c_null_ptr, nmax)
  if (rc /= ALPSG_OK) stop 2
  write(*, '(a,2i6)') 'NMAX', nmax
  om = [(0.5d0, 0.05d0), (0.5d0, -0.05d0), (0.5d0, -0.001d0), (3.3d0, -0.2d0)]
  do k = 1, 4
     det = alpsg_disp(h, om(k), rc)
     if (rc /= ALPSG_OK) stop 3
     write(*, '(a,4es25.16e3)') 'DISP', om(k), det
  enddo
  rp%secant_method = 2; rp%numiter = 50; rp%D_threshold = 1.d-15
  rp%D_prec = 1.d-5; rp%D_gap = 1.d-5; rp%D_tol = 1.d-7
  root = (0.5d0, -0.05d0)
  call alpsg_secant_osc(h, rp, root, rc)
  if (rc /= ALPSG_OK) stop 4
  write(*, '(a,2es25.16e3)') 'ROOT', root

  ! om_double_scan from kperp = kpar = 0.05 (scan_read's log steps)
  rc = alpsg_set_k(h, 0.05d0, 0.05d0, c_null_ptr, nmax)
  if (rc /= ALPSG_OK) stop 5
  rp%D_threshold = 1.d-14
  droots(1) = (0.048d0, -0.0001d0)
  call alpsg_secant_osc(h, rp, droots(1), rc)
  if (rc /= ALPSG_OK) stop 6
  s1 = alpsg_scan_params(4, 1, (log10(0.2d0) - log10(0.05d0)) / 3.d0, 0.d0, 3, 1)
  s2 = alpsg_scan_params(3, 1, (log10(0.2d0) - log10(0.05d0)) / 4.d0, 0.d0, 2, 2)
  rc = alpsg_om_double_scan(h, rp, s1, s2, 1_c_int, droots, 0_c_int, 0_c_int, kpo, kzo, rout, &
       written, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr)
  if (rc /= ALPSG_OK) then
     write(*, '(a,i0,2a)') 'alpsg_om_double_scan failed ', rc, ': ', trim(alpsg_error_text())
     stop 7
  endif
  do k = 1, nrow
     write(*, '(a,i4,4es25.16e3,i3)') 'DSCAN', k, kpo(k), kzo(k), rout(k), written(k)
  enddo
  rc = alpsg_destroy(h)

  ! two secant_osc solves on the root server, then its counters
  rc = alpsg_server_create(1.d9, 16_c_int, srv)
  if (rc /= ALPSG_OK) stop 8
  rc = alpsg_server_add_plasma(srv, d, pid)
  if (rc /= ALPSG_OK) stop 9
  g = [(0.29d0, -0.001d0), (0.5d0, -0.05d0)]
  tags = [11_c_long_long, 12_c_long_long]
  rc = alpsg_server_submit_batch(srv, 2_c_int, [pid, pid], [0.3d0, 0.3d0], [0.3d0, 0.3d0], g, rp, &
       tags)
  if (rc /= ALPSG_OK) stop 10
  nout = 0
  do while (nout < 2)
     rc = alpsg_server_poll(srv, 2_c_int - nout, 1_c_int, tags(nout + 1:), sroots(nout + 1:), &
          iters(nout + 1:), status(nout + 1:), lat(nout + 1:), k)
     if (rc /= ALPSG_OK) stop 11
     nout = nout + k
  enddo
  do k = 1, 2
     write(*, '(a,i4,2es25.16e3,2i4)') 'SERVED', tags(k), sroots(k), iters(k), status(k)
  enddo
  rc = alpsg_server_stats(srv, counts, times)
  if (rc /= ALPSG_OK) stop 12
  write(*, '(a,10i12)') 'STATS', counts
  rc = alpsg_server_destroy(srv)
end program alpsg_demo
