/*
 * orc_kpm.c -- operator kernels, Chebyshev moments, damping, summation,
 * Bessel time evolution and the quantity drivers, restated from the
 * reference's CPU build.  Test infrastructure only (see lsqt_oracle.h).
 *
 * Every per-site expression keeps the reference's operation order
 * (on-site term, then neighbours in list order, then the division by
 * energy_max), and this file is compiled with -ffp-contract=off, so each
 * state vector is bit-identical to the reference CPU build's.
 */
#define _DEFAULT_SOURCE
#include "lsqt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793
#define ORC_CHUNK 512 /* BLOCK_SIZE of src/vector.cu:22 */

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0)
    omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_ham_from_model(const orc_model* m, orc_ham* h)
{
  h->n = m->number_of_atoms;
  h->max_neighbor = m->max_neighbor;
  h->energy_max = m->energy_max;
  h->neighbor_number = m->neighbor_number;
  h->neighbor_list = m->neighbor_list;
  h->potential = m->potential;
  h->hopping_real = m->hopping_real;
  h->hopping_imag = m->hopping_imag;
  h->xx = m->xx;
}

/* src/hamiltonian.cu:188-221 */
void orc_apply_hamiltonian(const orc_ham* h, const double* ir, const double* ii, double* outr,
                           double* outi)
{
  const int K = h->max_neighbor;
#pragma omp parallel for schedule(static)
  for (int n = 0; n < h->n; ++n) {
    double tr = h->potential[n] * ir[n];
    double ti = h->potential[n] * ii[n];
    for (int m = 0; m < h->neighbor_number[n]; ++m) {
      size_t q = (size_t)n * K + m;
      int j = h->neighbor_list[q];
      double a = h->hopping_real[q], b = h->hopping_imag[q];
      double c = ir[j], d = ii[j];
      tr += a * c - b * d;
      ti += a * d + b * c;
    }
    tr /= h->energy_max;
    ti /= h->energy_max;
    outr[n] = tr;
    outi[n] = ti;
  }
}

/* src/hamiltonian.cu:273-304: [X, H~] */
void orc_apply_commutator(const orc_ham* h, const double* ir, const double* ii, double* outr,
                          double* outi)
{
  const int K = h->max_neighbor;
#pragma omp parallel for schedule(static)
  for (int n = 0; n < h->n; ++n) {
    double tr = 0.0, ti = 0.0;
    for (int m = 0; m < h->neighbor_number[n]; ++m) {
      size_t q = (size_t)n * K + m;
      int j = h->neighbor_list[q];
      double a = h->hopping_real[q], b = h->hopping_imag[q];
      double c = ir[j], d = ii[j];
      double x = h->xx[q];
      tr -= (a * c - b * d) * x;
      ti -= (a * d + b * c) * x;
    }
    outr[n] = tr / h->energy_max;
    outi[n] = ti / h->energy_max;
  }
}

/* src/hamiltonian.cu:354-383: V = i[H, X], not scaled by energy_max */
void orc_apply_current(const orc_ham* h, const double* ir, const double* ii, double* outr,
                       double* outi)
{
  const int K = h->max_neighbor;
#pragma omp parallel for schedule(static)
  for (int n = 0; n < h->n; ++n) {
    double tr = 0.0, ti = 0.0;
    for (int m = 0; m < h->neighbor_number[n]; ++m) {
      size_t q = (size_t)n * K + m;
      int j = h->neighbor_list[q];
      double a = h->hopping_real[q], b = h->hopping_imag[q];
      double c = ir[j], d = ii[j];
      tr += (a * c - b * d) * h->xx[q];
      ti += (a * d + b * c) * h->xx[q];
    }
    outr[n] = +ti;
    outi[n] = -tr;
  }
}

/* src/hamiltonian.cu:927-966: psi2 = 2 H~ psi1 - psi0 */
void orc_kernel_polynomial(const orc_ham* h, const double* s0r, const double* s0i,
                           const double* s1r, const double* s1i, double* s2r, double* s2i)
{
  const int K = h->max_neighbor;
#pragma omp parallel for schedule(static)
  for (int n = 0; n < h->n; ++n) {
    double tr = h->potential[n] * s1r[n];
    double ti = h->potential[n] * s1i[n];
    for (int m = 0; m < h->neighbor_number[n]; ++m) {
      size_t q = (size_t)n * K + m;
      int j = h->neighbor_list[q];
      double a = h->hopping_real[q], b = h->hopping_imag[q];
      double c = s1r[j], d = s1i[j];
      tr += a * c - b * d;
      ti += a * d + b * c;
    }
    tr /= h->energy_max;
    ti /= h->energy_max;
    tr = 2.0 * tr - s0r[n];
    ti = 2.0 * ti - s0i[n];
    s2r[n] = tr;
    s2i[n] = ti;
  }
}

/* src/hamiltonian.cu:425-443 */
void orc_chebyshev_01(int N, const double* s0r, const double* s0i, const double* s1r,
                      const double* s1i, double* sr, double* si, double b0, double b1,
                      int direction)
{
#pragma omp parallel for schedule(static)
  for (int n = 0; n < N; ++n) {
    double bessel_0 = b0;
    double bessel_1 = b1 * direction;
    sr[n] = bessel_0 * s0r[n] + bessel_1 * s1i[n];
    si[n] = bessel_0 * s0i[n] - bessel_1 * s1r[n];
  }
}

/* the four (-i)^m phase cases of src/hamiltonian.cu:570-590 */
static inline void accumulate_label(double* sr, double* si, double b, double tr, double ti,
                                    int label)
{
  switch (label) {
    case 1:
      *sr += b * tr;
      *si += b * ti;
      break;
    case 2:
      *sr -= b * tr;
      *si -= b * ti;
      break;
    case 3:
      *sr += b * ti;
      *si -= b * tr;
      break;
    case 4:
      *sr -= b * ti;
      *si += b * tr;
      break;
  }
}

/* src/hamiltonian.cu:531-595 */
void orc_chebyshev_2(const orc_ham* h, const double* s0r, const double* s0i, const double* s1r,
                     const double* s1i, double* s2r, double* s2i, double* sr, double* si,
                     double bessel_m, int label)
{
  const int K = h->max_neighbor;
#pragma omp parallel for schedule(static)
  for (int n = 0; n < h->n; ++n) {
    double tr = h->potential[n] * s1r[n];
    double ti = h->potential[n] * s1i[n];
    for (int m = 0; m < h->neighbor_number[n]; ++m) {
      size_t q = (size_t)n * K + m;
      int j = h->neighbor_list[q];
      double a = h->hopping_real[q], b = h->hopping_imag[q];
      double c = s1r[j], d = s1i[j];
      tr += a * c - b * d;
      ti += a * d + b * c;
    }
    tr /= h->energy_max;
    ti /= h->energy_max;
    tr = 2.0 * tr - s0r[n];
    ti = 2.0 * ti - s0i[n];
    accumulate_label(&sr[n], &si[n], bessel_m, tr, ti, label);
    s2r[n] = tr;
    s2i[n] = ti;
  }
}

/* src/hamiltonian.cu:635-648 */
void orc_chebyshev_1x(int N, const double* s1xr, const double* s1xi, double* sr, double* si,
                      double b1)
{
#pragma omp parallel for schedule(static)
  for (int n = 0; n < N; ++n) {
    sr[n] = +b1 * s1xi[n];
    si[n] = -b1 * s1xr[n];
  }
}

/* src/hamiltonian.cu:760-852 */
void orc_chebyshev_2x(const orc_ham* h, const double* s0r, const double* s0i,
                      const double* s0xr, const double* s0xi, const double* s1r,
                      const double* s1i, const double* s1xr, const double* s1xi, double* s2r,
                      double* s2i, double* s2xr, double* s2xi, double* sr, double* si,
                      double bessel_m, int label)
{
  const int K = h->max_neighbor;
#pragma omp parallel for schedule(static)
  for (int n = 0; n < h->n; ++n) {
    double tr = h->potential[n] * s1r[n];
    double ti = h->potential[n] * s1i[n];
    double txr = h->potential[n] * s1xr[n];
    double txi = h->potential[n] * s1xi[n];
    for (int m = 0; m < h->neighbor_number[n]; ++m) {
      size_t q = (size_t)n * K + m;
      int j = h->neighbor_list[q];
      double a = h->hopping_real[q], b = h->hopping_imag[q];
      double c = s1r[j], d = s1i[j];
      tr += a * c - b * d;
      ti += a * d + b * c;
      double cx = s1xr[j], dx = s1xi[j];
      txr += a * cx - b * dx;
      txi += a * dx + b * cx;
      double x = h->xx[q];
      txr -= (a * c - b * d) * x;
      txi -= (a * d + b * c) * x;
    }
    tr /= h->energy_max;
    ti /= h->energy_max;
    tr = 2.0 * tr - s0r[n];
    ti = 2.0 * ti - s0i[n];
    s2r[n] = tr;
    s2i[n] = ti;
    txr /= h->energy_max;
    txi /= h->energy_max;
    txr = 2.0 * txr - s0xr[n];
    txi = 2.0 * txi - s0xi[n];
    s2xr[n] = txr;
    s2xi[n] = txi;
    accumulate_label(&sr[n], &si[n], bessel_m, txr, txi, label);
  }
}

/* src/vector.cu:300-328 then :403-422: each 512-site chunk summed in site
 * order, then the chunk sums in chunk order. */
void orc_inner_product(int N, const double* pr, const double* pi, const double* Lr,
                       const double* Li, double* out_re, double* out_im)
{
  int nchunk = (N - 1) / ORC_CHUNK + 1;
  double* part = (double*)malloc(sizeof(double) * 2 * (size_t)nchunk);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nchunk; ++c) {
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < ORC_CHUNK; ++k) {
      int n = c * ORC_CHUNK + k;
      if (n < N) {
        double a = pr[n], b = pi[n], cc = Lr[n], d = Li[n];
        sr += (a * cc + b * d);
        si += (b * cc - a * d);
      }
    }
    part[2 * c] = sr;
    part[2 * c + 1] = si;
  }
  double sr = 0.0, si = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    sr += part[2 * c];
    si += part[2 * c + 1];
  }
  free(part);
  *out_re = sr;
  *out_im = si;
}

static double* vec_new(int n)
{
  return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
}

/* src/sigma.cu:33-61 */
void orc_find_moments(const orc_ham* h, int M, const double* Lr, const double* Li,
                      const double* Rr, const double* Ri, double* mu_re, double* mu_im)
{
  int N = h->n;
  double *s0r = vec_new(N), *s0i = vec_new(N);
  double *s1r = vec_new(N), *s1i = vec_new(N);
  double *s2r = vec_new(N), *s2i = vec_new(N);
  memcpy(s0r, Rr, sizeof(double) * (size_t)N);
  memcpy(s0i, Ri, sizeof(double) * (size_t)N);
  orc_inner_product(N, s0r, s0i, Lr, Li, &mu_re[0], &mu_im[0]);
  orc_apply_hamiltonian(h, s0r, s0i, s1r, s1i);
  orc_inner_product(N, s1r, s1i, Lr, Li, &mu_re[1], &mu_im[1]);
  for (int m = 2; m < M; ++m) {
    orc_kernel_polynomial(h, s0r, s0i, s1r, s1i, s2r, s2i);
    orc_inner_product(N, s2r, s2i, Lr, Li, &mu_re[m], &mu_im[m]);
    double* t;
    t = s0r; s0r = s1r; s1r = s2r; s2r = t;
    t = s0i; s0i = s1i; s1i = s2i; s2i = t;
  }
  free(s0r); free(s0i); free(s1r); free(s1i); free(s2r); free(s2i);
}

/* src/sigma.cu:76-86 */
void orc_apply_damping(int M, double* mu_re, double* mu_im)
{
  for (int k = 0; k < M; ++k) {
    double factor = 1.0 / (M + 1);
    double damping =
      (1 - k * factor) * cos(k * ORC_PI * factor) + sin(k * ORC_PI * factor) * factor / tan(ORC_PI * factor);
    mu_re[k] *= damping;
    mu_im[k] *= damping;
  }
}

/* src/sigma.cu:64-74 (compile-time LORENTZ; note k / M is integer division) */
void orc_apply_damping_lorentz(int M, double* mu_re, double* mu_im)
{
  double lambda = 4.0;
  double f1 = sinh(lambda);
  for (int k = 0; k < M; ++k) {
    double f2 = sinh(lambda * (1.0 - k / M)) / f1;
    mu_re[k] *= f2;
    mu_im[k] *= f2;
  }
}

/* src/sigma.cu:89-110; real part only */
void orc_chebyshev_summation(int M, const double* mu_re, int NE, const double* energy,
                             double energy_max, double volume, double* out)
{
  for (int e = 0; e < NE; ++e) {
    double x = energy[e] / energy_max;
    double c0 = 1.0, c1 = x, c2;
    double temp = mu_re[1] * c1;
    for (int k = 2; k < M; ++k) {
      c2 = 2.0 * x * c1 - c0;
      c0 = c1;
      c1 = c2;
      temp += mu_re[k] * c2;
    }
    temp *= 2.0;
    temp += mu_re[0];
    temp *= 2.0 / (ORC_PI * volume);
    temp /= sqrt(1.0 - x * x);
    out[e] = temp / energy_max;
  }
}

int orc_bessel_stop(double t)
{
  int m = 2;
  for (; m < 1000000; ++m) {
    double b = jn(m, t);
    if (b < 1.0e-15 && b > -1.0e-15)
      break;
  }
  return m;
}

#define SWAP3(a, b, c) \
  do {                 \
    double* _t = a;    \
    a = b;             \
    b = c;             \
    c = _t;            \
  } while (0)

/* src/sigma.cu:116-151 */
void orc_evolve(const orc_ham* h, int direction, double t, double* sr, double* si)
{
  int N = h->n;
  double *s0r = vec_new(N), *s0i = vec_new(N);
  double *s1r = vec_new(N), *s1i = vec_new(N);
  double *s2r = vec_new(N), *s2i = vec_new(N);
  memcpy(s0r, sr, sizeof(double) * (size_t)N);
  memcpy(s0i, si, sizeof(double) * (size_t)N);
  orc_apply_hamiltonian(h, sr, si, s1r, s1i);
  double b0 = j0(t);
  double b1 = 2.0 * j1(t);
  orc_chebyshev_01(N, s0r, s0i, s1r, s1i, sr, si, b0, b1, direction);
  for (int m = 2; m < 1000000; ++m) {
    double bm = jn(m, t);
    if (bm < 1.0e-15 && bm > -1.0e-15)
      break;
    bm *= 2.0;
    int label, r = m % 4;
    if (r == 0)
      label = 1;
    else if (r == 2)
      label = 2;
    else if ((r == 1 && direction == 1) || (r == 3 && direction == -1))
      label = 3;
    else
      label = 4;
    orc_chebyshev_2(h, s0r, s0i, s1r, s1i, s2r, s2i, sr, si, bm, label);
    SWAP3(s0r, s1r, s2r);
    SWAP3(s0i, s1i, s2i);
  }
  free(s0r); free(s0i); free(s1r); free(s1i); free(s2r); free(s2i);
}

/* src/sigma.cu:157-200 */
void orc_evolvex(const orc_ham* h, int direction, double t, double* sr, double* si)
{
  int N = h->n;
  double *s0r = vec_new(N), *s0i = vec_new(N), *s0xr = vec_new(N), *s0xi = vec_new(N);
  double *s1r = vec_new(N), *s1i = vec_new(N), *s1xr = vec_new(N), *s1xi = vec_new(N);
  double *s2r = vec_new(N), *s2i = vec_new(N), *s2xr = vec_new(N), *s2xi = vec_new(N);
  memcpy(s0r, sr, sizeof(double) * (size_t)N);
  memcpy(s0i, si, sizeof(double) * (size_t)N);
  orc_apply_hamiltonian(h, sr, si, s1r, s1i);
  orc_apply_commutator(h, sr, si, s1xr, s1xi);
  double b1 = 2.0 * j1(t);
  orc_chebyshev_1x(N, s1xr, s1xi, sr, si, b1);
  for (int m = 2; m <= 1000000; ++m) {
    double bm = jn(m, t);
    if (bm < 1.0e-15 && bm > -1.0e-15)
      break;
    bm *= 2.0;
    int label, r = m % 4;
    if (r == 1)
      label = 3;
    else if (r == 3)
      label = 4;
    else if ((r == 0 && direction == 1) || (r == 2 && direction == -1))
      label = 1;
    else
      label = 2;
    orc_chebyshev_2x(h, s0r, s0i, s0xr, s0xi, s1r, s1i, s1xr, s1xi, s2r, s2i, s2xr, s2xi, sr,
                     si, bm, label);
    SWAP3(s0r, s1r, s2r);
    SWAP3(s0i, s1i, s2i);
    SWAP3(s0xr, s1xr, s2xr);
    SWAP3(s0xi, s1xi, s2xi);
  }
  free(s0r); free(s0i); free(s0xr); free(s0xi);
  free(s1r); free(s1i); free(s1xr); free(s1xi);
  free(s2r); free(s2i); free(s2xr); free(s2xi);
}

/* one moment set -> one damped, summed row; optional raw moments */
static void moments_to_row(orc_model* m, const orc_ham* h, const double* Lr, const double* Li,
                           const double* Rr, const double* Ri, double* row, double* mu_out)
{
  int M = m->number_of_moments;
  double* mr = vec_new(M);
  double* mi = vec_new(M);
  orc_find_moments(h, M, Lr, Li, Rr, Ri, mr, mi);
  if (mu_out) {
    memcpy(mu_out, mr, sizeof(double) * (size_t)M);
    memcpy(mu_out + M, mi, sizeof(double) * (size_t)M);
  }
  orc_apply_damping(M, mr, mi);
  orc_chebyshev_summation(M, mr, m->number_of_energy_points, m->energy, m->energy_max,
                          m->volume, row);
  free(mr);
  free(mi);
}

#define MU_ROW(mu, r, M) ((mu) ? (mu) + (size_t)(r) * 2 * (M) : NULL)

/* src/sigma.cu:204-239 */
void orc_find_dos(orc_model* m, const orc_ham* h, const double* rr, const double* ri,
                  double* rows, double* mu)
{
  moments_to_row(m, h, rr, ri, rr, ri, rows, mu);
}

/* src/sigma.cu:243-274 */
void orc_find_vac0(orc_model* m, const orc_ham* h, const double* rr, const double* ri,
                   double* rows, double* mu)
{
  int N = h->n;
  double *sr = vec_new(N), *si = vec_new(N);
  orc_apply_current(h, rr, ri, sr, si);
  moments_to_row(m, h, sr, si, sr, si, rows, mu);
  free(sr);
  free(si);
}

/* src/sigma.cu:278-327 */
void orc_find_vac(orc_model* m, const orc_ham* h, const double* rr, const double* ri,
                  double* rows, double* mu)
{
  int N = h->n, NE = m->number_of_energy_points, M = m->number_of_moments;
  double *lr = vec_new(N), *li = vec_new(N);
  double *lcr = vec_new(N), *lci = vec_new(N);
  double *rgr = vec_new(N), *rgi = vec_new(N);
  memcpy(lr, rr, sizeof(double) * (size_t)N);
  memcpy(li, ri, sizeof(double) * (size_t)N);
  orc_apply_current(h, lr, li, rgr, rgi);
  for (int s = 0; s < m->number_of_steps_correlation; ++s) {
    orc_apply_current(h, lr, li, lcr, lci);
    moments_to_row(m, h, rgr, rgi, lcr, lci, rows + (size_t)s * NE, MU_ROW(mu, s, M));
    if (s < m->number_of_steps_correlation - 1) {
      double t = m->time_step[s] * m->energy_max;
      orc_evolve(h, -1, t, lr, li);
      orc_evolve(h, -1, t, rgr, rgi);
    }
  }
  free(lr); free(li); free(lcr); free(lci); free(rgr); free(rgi);
}

/* src/sigma.cu:331-391 */
void orc_find_msd(orc_model* m, const orc_ham* h, const double* rr, const double* ri,
                  double* rows, double* mu)
{
  int N = h->n, NE = m->number_of_energy_points, M = m->number_of_moments;
  double *sr = vec_new(N), *si = vec_new(N);
  double *xr = vec_new(N), *xi = vec_new(N);
  double *cr = vec_new(N), *ci = vec_new(N);
  memcpy(sr, rr, sizeof(double) * (size_t)N);
  memcpy(si, ri, sizeof(double) * (size_t)N);
  memcpy(xr, rr, sizeof(double) * (size_t)N);
  memcpy(xi, ri, sizeof(double) * (size_t)N);
  double t = m->time_step[0] * m->energy_max;
  orc_evolve(h, 1, t, sr, si);
  orc_evolvex(h, 1, t, xr, xi);
  for (int s = 0; s < m->number_of_steps_correlation; ++s) {
    moments_to_row(m, h, xr, xi, xr, xi, rows + (size_t)s * NE, MU_ROW(mu, s, M));
    if (s < m->number_of_steps_correlation - 1) {
      t = m->time_step[s + 1] * m->energy_max;
      memcpy(cr, sr, sizeof(double) * (size_t)N);
      memcpy(ci, si, sizeof(double) * (size_t)N);
      orc_evolvex(h, 1, t, cr, ci);
      orc_evolve(h, 1, t, xr, xi);
      for (int n = 0; n < N; ++n) {
        xr[n] += cr[n];
        xi[n] += ci[n];
      }
      orc_evolve(h, 1, t, sr, si);
    }
  }
  free(sr); free(si); free(xr); free(xi); free(cr); free(ci);
}

/* src/sigma.cu:395-442 */
void orc_find_spin(orc_model* m, const orc_ham* h, const double* rr, const double* ri,
                   double* rows, double* mu)
{
  int N = h->n, NE = m->number_of_energy_points, M = m->number_of_moments;
  double *sr = vec_new(N), *si = vec_new(N);
  double *zr = vec_new(N), *zi = vec_new(N);
  memcpy(sr, rr, sizeof(double) * (size_t)N);
  memcpy(si, ri, sizeof(double) * (size_t)N);
  for (int s = 0; s < m->number_of_steps_correlation; ++s) {
    for (int n = 0; n < N; ++n) { /* src/vector.cu:174-185 */
      if (n % 2 == 0) {
        zr[n] = sr[n];
        zi[n] = si[n];
      } else {
        zr[n] = -sr[n];
        zi[n] = -si[n];
      }
    }
    moments_to_row(m, h, sr, si, zr, zi, rows + (size_t)s * NE, MU_ROW(mu, s, M));
    if (s < m->number_of_steps_correlation - 1) {
      double t = m->time_step[s] * m->energy_max;
      orc_evolve(h, 1, t, sr, si);
    }
  }
  free(sr); free(si); free(zr); free(zi);
}

int orc_format_row(const double* v, int n, char* buf, int cap)
{
  int off = 0;
  for (int i = 0; i < n; ++i) {
    int w = snprintf(buf + off, (size_t)(cap - off), "%g ", v[i]);
    if (w < 0 || w >= cap - off)
      return -1;
    off += w;
  }
  if (off + 1 >= cap)
    return -1;
  buf[off++] = '\n';
  buf[off] = 0;
  return off;
}

static int append_rows(const char* dir, const char* name, const double* rows, int nrows,
                       int ncols, char* err, int err_len)
{
  char path[4200];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE* f = fopen(path, "a");
  if (!f) {
    snprintf(err, (size_t)err_len, "Error: cannot open %s", path);
    return 1;
  }
  int cap = 32 * ncols + 16;
  char* buf = (char*)malloc((size_t)cap);
  for (int r = 0; r < nrows; ++r) {
    orc_format_row(rows + (size_t)r * ncols, ncols, buf, cap);
    fputs(buf, f);
  }
  free(buf);
  fclose(f);
  return 0;
}

/* src/lsqt.cu:135-152 */
int orc_lsqt(const char* dir, char* err, int err_len)
{
  orc_model m;
  if (orc_model_load(dir, &m)) {
    snprintf(err, (size_t)err_len, "%s", m.error);
    orc_model_free(&m);
    return 1;
  }
  orc_ham h;
  orc_ham_from_model(&m, &h);
  int N = m.number_of_atoms, NE = m.number_of_energy_points, Nt = m.number_of_steps_correlation;
  double *rr = vec_new(N), *ri = vec_new(N);
  double* rows = vec_new(NE * (Nt > 1 ? Nt : 1));
  int rc = 0;
  for (int i = 0; i < m.number_of_random_vectors && !rc; ++i) {
    orc_initialize_state(&m, -1, rr, ri);
    orc_find_dos(&m, &h, rr, ri, rows, NULL);
    rc |= append_rows(dir, "dos.out", rows, 1, NE, err, err_len);
    if (m.calculate_vac0) {
      orc_find_vac0(&m, &h, rr, ri, rows, NULL);
      rc |= append_rows(dir, "vac0.out", rows, 1, NE, err, err_len);
    }
    if (m.calculate_vac) {
      orc_find_vac(&m, &h, rr, ri, rows, NULL);
      rc |= append_rows(dir, "vac.out", rows, Nt, NE, err, err_len);
    }
    if (m.calculate_msd) {
      orc_find_msd(&m, &h, rr, ri, rows, NULL);
      rc |= append_rows(dir, "msd.out", rows, Nt, NE, err, err_len);
    }
    if (m.calculate_spin) {
      orc_find_spin(&m, &h, rr, ri, rows, NULL);
      rc |= append_rows(dir, "S.out", rows, Nt, NE, err, err_len);
    }
  }
  if (m.calculate_ldos && !rc) {
    for (int i = 0; i < m.number_of_local_orbitals; ++i) {
      orc_initialize_state(&m, m.local_orbitals[i], rr, ri);
      orc_find_dos(&m, &h, rr, ri, rows, NULL);
      rc |= append_rows(dir, "ldos.out", rows, 1, NE, err, err_len);
    }
  }
  free(rr);
  free(ri);
  free(rows);
  orc_model_free(&m);
  return rc;
}
