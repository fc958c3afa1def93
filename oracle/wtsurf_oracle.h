/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.  Never linked into, loaded by or
 * called from the product path (wannier_tools_amd/ and libwtsurf.so).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker.
 *
 * A from-scratch FP64 C restatement of the WannierTools surface-state path:
 *   wto_ham_qlayer2qlayer  <- src/ham_qlayer2qlayer.f90:1-102
 *   wto_inv                <- src/inverse.f90:1-58   (zgesv, identity RHS)
 *   wto_mat_mul            <- src/mat_mul.f90:3-27   (zgemm 'N','N')
 *   wto_surfgreen_1985     <- src/surfgreen.f90:7-137
 *   wto_spectral           <- src/surfstat.f90:123-160 / src/fermiarc.f90:206-248
 *   wto_spectral_spin      <- src/surfstat.f90:102-115,163-192 /
 *                             src/fermiarc.f90:85-101,250-328
 *   wto_kplane_autocorrelation <- src/fermiarc.f90:496-560 (QPI jdos / jsdos)
 *   wto_kpath_jdos         <- src/surfstat.f90:891-907 (surfstat_jdos)
 *
 * Conventions (all as in the reference): complex matrices are column-major,
 * interleaved (re, im) doubles, leading dimension = order; 1-based orbital
 * index lists; H(R) already divided by ndegen (src/readHmnR.f90:720).
 *
 * Pinning (each function is checked by tests/test_oracle.py):
 *   wto_inv, wto_mat_mul      -- numpy inv / matmul, and the reference's own
 *                                inv / mat_mul (oracle/_ref) where built.
 *   wto_ham_qlayer2qlayer     -- reference golden vectors (tests/golden/) and
 *                                a hand-evaluated single-R closed form.
 *   wto_surfgreen_1985        -- reference golden vectors, the 1-D chain
 *                                closed form, block-diagonal chains, and a
 *                                finite-slab brute force.
 *   wto_spectral              -- reference golden traces and the diagonal
 *                                sums of wto_surfgreen_1985.
 *   wto_spectral_spin         -- the reference program's spindos.dat_l/_r
 *                                (tests/golden/files_*), and spin-diagonal
 *                                models where s_z must equal the up minus
 *                                down spectral weight and s_x = s_y = 0.
 *   wto_kplane_autocorrelation -- the reference program's arc.jdat_* /
 *                                arc.jsdat_*, and an FFT cyclic correlation
 *                                of the folded plane.
 *   wto_kpath_jdos            -- the reference program's dos.jdat_*, and
 *                                numpy's full linear correlation.
 */
#ifndef WTSURF_ORACLE_H
#define WTSURF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fourier sum H00(k), H01(k).  irvec is [nr][3] (R1,R2,R3), hr is
 * [nr][nw][nw] column-major per R.  h00, h01 are N x N, N = nw*np. */
int wto_ham_qlayer2qlayer(int nw, int np, int nr, int ijmax, const int *irvec,
                          const double *hr, const double *k, double *h00,
                          double *h01);

/* In-place inverse by LU with partial pivoting (largest |re|+|im|, the
 * LAPACK izamax rule) and triangular solves against the identity.  Returns
 * the zgesv-style info (0, or the 1-based index of an exactly zero pivot). */
int wto_inv(int n, double *a);

/* C = A * B. */
void wto_mat_mul(int n, const double *a, const double *b, double *c);

/* Lopez-Sancho decimation; writes the full GLL, GRR, GB.  *iters receives the
 * number of decimation iterations executed (the loop body count). */
int wto_surfgreen_1985(int n, double omega, double eta, const double *h00,
                       const double *h01, double *gll, double *grr, double *gb,
                       int itermax, double accuracy, int *iters);

/* Batched traces over an nk x nomega grid (H00/H01 built once per k, as the
 * reference does).  Outputs are [nomega][nk] (ik fastest), like the
 * reference's dos_l(knv2, omeganum).  Every trace is divided by divisor (1 for
 * surfstat, pi = 3.14159265358979 for SurfaceDOSkk).  iters may be NULL. */
int wto_spectral(int nw, int np, int nr, int ijmax, const int *irvec,
                 const double *hr, int nk, const double *kpts, int nomega,
                 const double *omega, double eta, int ntop, const int *top,
                 int nbot, const int *bot, int itermax, double accuracy,
                 double divisor, int nthreads, double *dos_l, double *dos_r,
                 double *dos_bulk, int *iters);

/* wto_spectral plus the spin-resolved traces (src/surfstat.f90:102-115,
 * 163-192; src/fermiarc.f90:85-101, 250-328): s_c = -sum Im (G sigma_c)(io,io)
 * / divisor with G sigma_c formed as a full product.  Left: GLL over top.
 * Right: GRR over top (spin_mode 1, surfstat) or over bot offset by N - nw
 * (spin_mode 2, SurfaceDOSkk).  spin_l, spin_r are [3][nomega][nk]. */
int wto_spectral_spin(int nw, int np, int nr, int ijmax, const int *irvec,
                      const double *hr, int nk, const double *kpts, int nomega,
                      const double *omega, double eta, int ntop, const int *top,
                      int nbot, const int *bot, int itermax, double accuracy,
                      double divisor, int nthreads, double *dos_l, double *dos_r,
                      double *dos_bulk, int *iters, int spin_mode, double *spin_l,
                      double *spin_r);

/* QPI autocorrelation over a k-plane (src/fermiarc.f90:496-560):
 * out(q) = sum_{k} sum_f a_f(k) a_f((k + q) mod (m1, m2)) with the
 * reference's 1-based index arithmetic; m1 = nkx - 1, m2 = nky - 1 for jdos,
 * m1 = Nk1 - 1, m2 = Nk2 - 1 for jsdos.  fields [nfield][nky][nkx], out
 * [nky][nkx], ik1 fastest. */
int wto_kplane_autocorrelation(int nkx, int nky, int m1, int m2, int nfield,
                               const double *fields, double *out);

/* k-path joint density of states of surfstat_jdos (src/surfstat.f90:891-907):
 * out_f(iq, j) = sum_{ik1} a_f(ik1, j) a_f(ik1 + iq - (nk-1)/2, j) over the
 * reference's index window.  fields, out [nfield][nomega][nk], k fastest. */
int wto_kpath_jdos(int nk, int nomega, int nfield, const double *fields, double *out);

#ifdef __cplusplus
}
#endif
#endif
