!! ISO_C_BINDING interface of the gfx950 dispersion engine (include/alpsg.h)
!! for the ALPS Fortran host.  The module mirrors the reference's call
!! surface: alpsg_disp(om) stands in for disp(om) (ALPS_fns.f90:252),
!! alpsg_secant_osc for secant_osc (ALPS_fns.f90:1919), alpsg_map for
!! map_search (ALPS_fns.f90:3595).  Arrays are passed exactly as alps_var
!! holds them (column-major, species index fastest).
module alpsg
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ALPSG_OK = 0, ALPSG_E_ARG = 1, &
       ALPSG_E_UNSUPPORTED = 2, ALPSG_E_DEVICE = 3, ALPSG_E_STATE = 4, ALPSG_E_NOTFOUND = 5, &
       ALPSG_MAXITER = 6

  type, bind(C), public :: alpsg_plasma_desc
     integer(c_int) :: nspec, nperp, npar
     type(c_ptr) :: pp = c_null_ptr, f0 = c_null_ptr, df0 = c_null_ptr
     type(c_ptr) :: ns = c_null_ptr, qs = c_null_ptr, ms = c_null_ptr
     real(c_double) :: vA
     integer(c_int) :: kperp_norm, positions_principal, n_resonance_interval
     real(c_double) :: Tlim, Bessel_zero
     type(c_ptr) :: usebM = c_null_ptr, relativistic = c_null_ptr
     type(c_ptr) :: ACmethod = c_null_ptr, n_fits = c_null_ptr
     integer(c_int) :: max_fits
     type(c_ptr) :: fit_type = c_null_ptr, perp_correction = c_null_ptr, param_fit = c_null_ptr
     integer(c_int) :: param_fit_nperp
     type(c_ptr) :: poly_kind = c_null_ptr, poly_order = c_null_ptr, logfit = c_null_ptr
     type(c_ptr) :: poly_log_max = c_null_ptr, poly_fit_coeffs = c_null_ptr
     integer(c_int) :: poly_max_order
     type(c_ptr) :: bM_nmax = c_null_ptr, bM_Bessel_zero = c_null_ptr, bM_beta = c_null_ptr
     type(c_ptr) :: bM_alpha = c_null_ptr, bM_pdrift = c_null_ptr
     integer(c_int) :: ngamma = 0, npparbar = 0
     type(c_ptr) :: gamma_rel = c_null_ptr, pparbar_rel = c_null_ptr
     type(c_ptr) :: f0_rel = c_null_ptr, df0_rel = c_null_ptr
     type(c_ptr) :: current_int = c_null_ptr
  end type alpsg_plasma_desc

  !! fit controls (&system: maxsteps_fit, lambda_initial_fit, lambdafac_fit, epsilon_fit)
  type, bind(C), public :: alpsg_fit_params
     integer(c_int) :: maxsteps_fit = 500
     real(c_double) :: lambda_initial_fit = 1.d0, lambdafac_fit = 10.d0, epsilon_fit = 1.d-8
  end type alpsg_fit_params

  type, bind(C), public :: alpsg_root_params
     integer(c_int) :: secant_method, numiter
     real(c_double) :: D_threshold, D_prec, D_gap, D_tol
  end type alpsg_root_params

  !! one &scan_input scan (scan_type, swlog, diff/diff2 as scan_read computes
  !! them, ALPS_io.f90:482-548, ns, nres)
  type, bind(C), public :: alpsg_scan_params
     integer(c_int) :: scan_type, log_scan
     real(c_double) :: diff, diff2
     integer(c_int) :: n_out, n_res
  end type alpsg_scan_params

  public :: alpsg_create, alpsg_destroy, alpsg_set_k, alpsg_disp_batch, alpsg_disp_plain
  public :: alpsg_secant_batch, alpsg_map_search, alpsg_om_scan, alpsg_get_nmax
  public :: alpsg_eval_count, alpsg_error_text, alpsg_disp, alpsg_secant_osc
  public :: alpsg_autotune, alpsg_set_shape, alpsg_get_shape, alpsg_calc_eigen
  public :: alpsg_server_create, alpsg_server_add_plasma, alpsg_server_remove_plasma
  public :: alpsg_server_submit_k, alpsg_server_submit_batch, alpsg_server_poll
  public :: alpsg_server_destroy, alpsg_determine_param_fit
  public :: alpsg_om_double_scan, alpsg_server_stats, alpsg_last_scan_steps, alpsg_set_plain_roots
  public :: alpsg_map_search_refine

  interface
     function alpsg_create(desc, h) bind(C, name='alpsg_create') result(rc)
       import :: c_int, c_ptr, alpsg_plasma_desc
       type(alpsg_plasma_desc), intent(in) :: desc
       type(c_ptr), intent(out) :: h
       integer(c_int) :: rc
     end function alpsg_create

     function alpsg_destroy(h) bind(C, name='alpsg_destroy') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int) :: rc
     end function alpsg_destroy

     function alpsg_set_k(h, kperp, kpar, nmax_in, nmax_out) bind(C, name='alpsg_set_k') result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: kperp, kpar
       type(c_ptr), value :: nmax_in
       integer(c_int), intent(out) :: nmax_out(*)
       integer(c_int) :: rc
     end function alpsg_set_k

     function alpsg_disp_batch(h, n, om, det, chi0, chi0_low, wave) &
          bind(C, name='alpsg_disp_batch') result(rc)
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: n
       complex(c_double_complex), intent(in) :: om(*)
       complex(c_double_complex), intent(out) :: det(*)
       type(c_ptr), value :: chi0, chi0_low, wave
       integer(c_int) :: rc
     end function alpsg_disp_batch

     function alpsg_disp_plain(h, n, om, det) bind(C, name='alpsg_disp_plain') result(rc)
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: n
       complex(c_double_complex), intent(in) :: om(*)
       complex(c_double_complex), intent(out) :: det(*)
       integer(c_int) :: rc
     end function alpsg_disp_plain

     function alpsg_secant_batch(h, rp, nroots, om, D_final, iters) &
          bind(C, name='alpsg_secant_batch') result(rc)
       import :: c_int, c_ptr, c_double_complex, alpsg_root_params
       type(c_ptr), value :: h
       type(alpsg_root_params), intent(in) :: rp
       integer(c_int), value :: nroots
       complex(c_double_complex), intent(inout) :: om(*)
       complex(c_double_complex), intent(out) :: D_final(*)
       integer(c_int), intent(out) :: iters(*)
       integer(c_int) :: rc
     end function alpsg_secant_batch

     function alpsg_map_search(h, omi, omf, gami, gamf, nr, ni, loggridw, loggridg, om_out, &
          cal_out, val_out, numroots, iroots, nfound) bind(C, name='alpsg_map_search') result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: omi, omf, gami, gamf
       integer(c_int), value :: nr, ni, loggridw, loggridg, numroots
       type(c_ptr), value :: om_out, cal_out, val_out
       integer(c_int), intent(out) :: iroots(2, *)
       integer(c_int), intent(out) :: nfound
       integer(c_int) :: rc
     end function alpsg_map_search

     ! map_search with refine_guess on its minima (ALPS_fns.f90:3595-3788,
     ! 3793-3856): roots, D_final, iters of the first min(nfound, numroots)
     ! minima
     function alpsg_map_search_refine(h, rp, omi, omf, gami, gamf, nr, ni, loggridw, loggridg, &
          om_out, cal_out, val_out, numroots, iroots, nfound, roots, D_final, iters) &
          bind(C, name='alpsg_map_search_refine') result(rc)
       import :: c_int, c_ptr, c_double, c_double_complex, alpsg_root_params
       type(c_ptr), value :: h
       type(alpsg_root_params), intent(in) :: rp
       real(c_double), value :: omi, omf, gami, gamf
       integer(c_int), value :: nr, ni, loggridw, loggridg, numroots
       type(c_ptr), value :: om_out, cal_out, val_out
       integer(c_int), intent(out) :: iroots(2, *)
       integer(c_int), intent(out) :: nfound
       complex(c_double_complex), intent(out) :: roots(*)
       complex(c_double_complex), intent(out) :: D_final(*)
       integer(c_int), intent(out) :: iters(*)
       integer(c_int) :: rc
     end function alpsg_map_search_refine

     function alpsg_om_scan(h, rp, scan_type, log_scan, diff, diff2, n_out, n_res, nroots, roots, &
          kperp_out, kpar_out, roots_out) bind(C, name='alpsg_om_scan') result(rc)
       import :: c_int, c_ptr, c_double, c_double_complex, alpsg_root_params
       type(c_ptr), value :: h
       type(alpsg_root_params), intent(in) :: rp
       integer(c_int), value :: scan_type, log_scan, n_out, n_res, nroots
       real(c_double), value :: diff, diff2
       complex(c_double_complex), intent(inout) :: roots(*)
       real(c_double), intent(out) :: kperp_out(*), kpar_out(*)
       complex(c_double_complex), intent(out) :: roots_out(*)
       integer(c_int) :: rc
     end function alpsg_om_scan

     !! om_double_scan (ALPS_fns.f90:2900-3591), called in its place at
     !! ALPS.f90:126; the eigen/heat arrays may be c_null_ptr when not asked for
     function alpsg_om_double_scan(h, rp, s1, s2, nroots, roots, eigen, heat, kperp_out, &
          kpar_out, roots_out, written, ef, bf, Us, ds, Ps, Ps_split, W_EM) &
          bind(C, name='alpsg_om_double_scan') result(rc)
       import :: c_int, c_ptr, c_double, c_double_complex, alpsg_root_params, alpsg_scan_params
       type(c_ptr), value :: h
       type(alpsg_root_params), intent(in) :: rp
       type(alpsg_scan_params), intent(in) :: s1, s2
       integer(c_int), value :: nroots, eigen, heat
       complex(c_double_complex), intent(inout) :: roots(*)
       real(c_double), intent(out) :: kperp_out(*), kpar_out(*)
       complex(c_double_complex), intent(out) :: roots_out(*)
       integer(c_int), intent(out) :: written(*)
       type(c_ptr), value :: ef, bf, Us, ds, Ps, Ps_split, W_EM
       integer(c_int) :: rc
     end function alpsg_om_double_scan

     function alpsg_last_scan_steps(h, nsteps, nroots, ms_out, iters_out) &
          bind(C, name='alpsg_last_scan_steps') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nsteps, nroots
       type(c_ptr), value :: ms_out, iters_out
       integer(c_int) :: rc
     end function alpsg_last_scan_steps

     function alpsg_set_plain_roots(h, plain) bind(C, name='alpsg_set_plain_roots') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: plain
       integer(c_int) :: rc
     end function alpsg_set_plain_roots

     function alpsg_get_nmax(h, nmax_out) bind(C, name='alpsg_get_nmax') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: nmax_out(*)
       integer(c_int) :: rc
     end function alpsg_get_nmax

     function alpsg_eval_count(h) bind(C, name='alpsg_eval_count') result(n)
       import :: c_ptr, c_long_long
       type(c_ptr), value :: h
       integer(c_long_long) :: n
     end function alpsg_eval_count

     function alpsg_calc_eigen(h, n, om, eigen, heat, ef, bf, Us, ds, Ps, Ps_split, W_EM) &
          bind(C, name='alpsg_calc_eigen') result(rc)
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: n, eigen, heat
       complex(c_double_complex), intent(in) :: om(*)
       type(c_ptr), value :: ef, bf, Us, ds, Ps, Ps_split, W_EM
       integer(c_int) :: rc
     end function alpsg_calc_eigen

     function alpsg_autotune(h, n, om, shape_out) bind(C, name='alpsg_autotune') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: n
       type(c_ptr), value :: om
       integer(c_int), intent(out) :: shape_out(5)
       integer(c_int) :: rc
     end function alpsg_autotune

     function alpsg_set_shape(h, n, shape) bind(C, name='alpsg_set_shape') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: n
       integer(c_int), intent(in) :: shape(5)
       integer(c_int) :: rc
     end function alpsg_set_shape

     function alpsg_get_shape(h, n, shape_out) bind(C, name='alpsg_get_shape') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: n
       integer(c_int), intent(out) :: shape_out(5)
       integer(c_int) :: rc
     end function alpsg_get_shape

     !! determine_param_fit (ALPS_analyt.f90:365) on the device; the outputs
     !! may be the description's own arrays (param_fit, fit_type, ...)
     function alpsg_determine_param_fit(desc, fp, param_fit, fit_type, perp_correction, &
          poly_fit_coeffs, quality) bind(C, name='alpsg_determine_param_fit') result(rc)
       import :: c_int, c_ptr, c_double, alpsg_plasma_desc, alpsg_fit_params
       type(alpsg_plasma_desc), intent(in) :: desc
       type(alpsg_fit_params), intent(in) :: fp
       type(c_ptr), value :: param_fit, fit_type, perp_correction, poly_fit_coeffs
       real(c_double), intent(out) :: quality
       integer(c_int) :: rc
     end function alpsg_determine_param_fit

     !! root server: queue secant_osc solves of any plasma and k, collect roots
     function alpsg_server_create(hbm_bytes, max_slots, s) &
          bind(C, name='alpsg_server_create') result(rc)
       import :: c_int, c_ptr, c_double
       real(c_double), value :: hbm_bytes
       integer(c_int), value :: max_slots
       type(c_ptr), intent(out) :: s
       integer(c_int) :: rc
     end function alpsg_server_create

     function alpsg_server_add_plasma(s, desc, plasma_id) &
          bind(C, name='alpsg_server_add_plasma') result(rc)
       import :: c_int, c_ptr, alpsg_plasma_desc
       type(c_ptr), value :: s
       type(alpsg_plasma_desc), intent(in) :: desc
       integer(c_int), intent(out) :: plasma_id
       integer(c_int) :: rc
     end function alpsg_server_add_plasma

     function alpsg_server_remove_plasma(s, plasma_id) &
          bind(C, name='alpsg_server_remove_plasma') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: s
       integer(c_int), value :: plasma_id
       integer(c_int) :: rc
     end function alpsg_server_remove_plasma

     function alpsg_server_submit_k(s, plasma_id, kperp, kpar, om_guess, rp, tag) &
          bind(C, name='alpsg_server_submit_k') result(rc)
       import :: c_int, c_ptr, c_double, c_double_complex, c_long_long, alpsg_root_params
       type(c_ptr), value :: s
       integer(c_int), value :: plasma_id
       real(c_double), value :: kperp, kpar
       complex(c_double_complex), intent(in) :: om_guess
       type(alpsg_root_params), intent(in) :: rp
       integer(c_long_long), value :: tag
       integer(c_int) :: rc
     end function alpsg_server_submit_k

     function alpsg_server_submit_batch(s, n, plasma_ids, kperp, kpar, om_guess, rp, tags) &
          bind(C, name='alpsg_server_submit_batch') result(rc)
       import :: c_int, c_ptr, c_double, c_double_complex, c_long_long, alpsg_root_params
       type(c_ptr), value :: s
       integer(c_int), value :: n
       integer(c_int), intent(in) :: plasma_ids(*)
       real(c_double), intent(in) :: kperp(*), kpar(*)
       complex(c_double_complex), intent(in) :: om_guess(*)
       type(alpsg_root_params), intent(in) :: rp
       integer(c_long_long), intent(in) :: tags(*)
       integer(c_int) :: rc
     end function alpsg_server_submit_batch

     function alpsg_server_poll(s, max_results, wait, tags, roots, iters, status, latency_ms, &
          n_out) bind(C, name='alpsg_server_poll') result(rc)
       import :: c_int, c_ptr, c_double, c_double_complex, c_long_long
       type(c_ptr), value :: s
       integer(c_int), value :: max_results, wait
       integer(c_long_long), intent(out) :: tags(*)
       complex(c_double_complex), intent(out) :: roots(*)
       integer(c_int), intent(out) :: iters(*), status(*)
       real(c_double), intent(out) :: latency_ms(*)
       integer(c_int), intent(out) :: n_out
       integer(c_int) :: rc
     end function alpsg_server_poll

     !! counts(10): submitted, finished, passes, disp evaluations, admitted,
     !! waiting, arena bytes, arena bytes in use, plasma uploads, evictions;
     !! times_ms(4): passes, admission, launch enqueue, device waits
     function alpsg_server_stats(s, counts, times_ms) bind(C, name='alpsg_server_stats') result(rc)
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: s
       integer(c_long_long), intent(out) :: counts(10)
       real(c_double), intent(out) :: times_ms(4)
       integer(c_int) :: rc
     end function alpsg_server_stats

     function alpsg_server_destroy(s) bind(C, name='alpsg_server_destroy') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: s
       integer(c_int) :: rc
     end function alpsg_server_destroy

     function alpsg_last_error_c() bind(C, name='alpsg_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function alpsg_last_error_c
  end interface

contains

  !> Text of the last error (alpsg_last_error) as a Fortran string.
  function alpsg_error_text() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = alpsg_last_error_c()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    enddo
  end function alpsg_error_text

  !> det D at one frequency: the drop-in body of disp(om) (ALPS_fns.f90:252).
  !! chi0, chi0_low and wave (alps_var side effects) are filled when present.
  function alpsg_disp(h, om, rc, chi0, chi0_low, wave) result(det)
    type(c_ptr), intent(in) :: h
    complex(c_double_complex), intent(in) :: om
    integer(c_int), intent(out) :: rc
    complex(c_double_complex), intent(out), target, optional, contiguous :: chi0(:,:,:), &
         chi0_low(:,:,:,:), wave(:,:)
    complex(c_double_complex) :: det
    complex(c_double_complex) :: o(1), d(1)
    type(c_ptr) :: pc, pl, pw
    pc = c_null_ptr; pl = c_null_ptr; pw = c_null_ptr
    if (present(chi0)) pc = c_loc(chi0)
    if (present(chi0_low)) pl = c_loc(chi0_low)
    if (present(wave)) pw = c_loc(wave)
    o(1) = om
    rc = alpsg_disp_batch(h, 1_c_int, o, d, pc, pl, pw)
    det = d(1)
  end function alpsg_disp

  !> secant_osc(om, in) (ALPS_fns.f90:1919) through the batched device solver.
  subroutine alpsg_secant_osc(h, rp, om, rc)
    type(c_ptr), intent(in) :: h
    type(alpsg_root_params), intent(in) :: rp
    complex(c_double_complex), intent(inout) :: om
    integer(c_int), intent(out) :: rc
    complex(c_double_complex) :: o(1), d(1)
    integer(c_int) :: it(1)
    o(1) = om
    rc = alpsg_secant_batch(h, rp, 1_c_int, o, d, it)
    om = o(1)
  end subroutine alpsg_secant_osc

end module alpsg
