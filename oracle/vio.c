/*
 * vio.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see vio.h). Compiled with
 * -ffp-contract=off so that every multiply and add rounds separately, as Julia does
 * without @fastmath/muladd (the reference's arithmetic), and so that the HIP kernels,
 * built with the same flag, can be compared with it bit for bit.
 */
#include "vio.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Staggered 8th-order first-derivative weights at offsets +-1/2 .. +-7/2, i.e. the
 * solution of _stencil(x, 0, 1) (src/common/Stencil.jl:9-14) for x = -7/2:1:7/2 folded
 * into the half stencil that diff-funcs.jl:3-32 applies. */
void vio_fd_coeffs(double h, double out[4]) {
  out[0] = (1225.0 / 1024.0) / h;
  out[1] = (-245.0 / 3072.0) / h;
  out[2] = (49.0 / 5120.0) / h;
  out[3] = (-5.0 / 7168.0) / h;
}

/* Staggered first-derivative weights of order 2K (Stencil(order; stagger), Stencil.jl:280-320:
 * _stencil at +-1/2 .. +-(2K-1)/2), by the closed form of the staggered-grid weights
 *   c_l = (-1)^(l+1) / (2l-1) * prod_{i != l} (2i-1)^2 / |(2l-1)^2 - (2i-1)^2|,  l = 1..K
 * (exact integers, one rounding by the division, then / h; unused entries 0). */
int vio_fd_weights(double h, int order, double out[4]) {
  if (order < 2 || order > 8 || order % 2 || !(h > 0)) return -1;
  const int K = order / 2;
  for (int l = 0; l < 4; ++l) out[l] = 0.0;
  for (int l = 1; l <= K; ++l) {
    long long num = 1, den = 2 * l - 1;
    for (int i = 1; i <= K; ++i) {
      if (i == l) continue;
      const long long a = (2 * i - 1) * (2 * i - 1), b = (2 * l - 1) * (2 * l - 1) - a;
      num *= a;
      den *= b < 0 ? -b : b;
    }
    out[l - 1] = ((l % 2 ? 1.0 : -1.0) * (double)num / (double)den) / h;
  }
  return 0;
}

/* CPMLCoefficients (src/sciml_solver/CPML.jl:109-189) along one axis, in the working
 * precision F of the arrays, as the reference computes them:
 *   lo strip i in [0,T):       p = (T - i - 0.5)/T                      (CPML.jl:124)
 *   hi strip i in [n-T, n):    p = (i - (n-T) + 0.5)/T                  (CPML.jl:134)
 *   d = F(d0 p^power), alpha = F(alpha_max (1-p)), kappa = F(1 + (kappa_max-1) p), the
 *   profiles evaluated in double and stored in F                        (CPML.jl:91-161)
 *   where d > 0 or alpha > 0:
 *     b = F(exp(-(d/kappa + alpha) dt)), d/kappa + alpha summed in F, the product with
 *     the double dt and the exponential in double                       (CPML.jl:172-174)
 *     a = d/(kappa (d + kappa alpha)) (b - 1) in F, where |d| > eps(F)  (CPML.jl:175-177)
 * vio_cpml_axis gives F = double, vio_cpml_axis_f32 F = float. */
#define VIO_CPML_AXIS(NAME_, F, EPS, ABS)                                                     \
  void NAME_(int n, int T, double d0, double alpha_max, double kappa_max, double power,       \
             double dt, int lo, int hi, F* a, F* b, F* kappa, F* dd, F* al) {                 \
    for (int i = 0; i < n; ++i) {                                                          \
      double p = -1.0;                                                                     \
      if (lo && i < T) p = (T - i - 0.5) / T;                                              \
      if (hi && i >= n - T) p = (i - (n - T) + 0.5) / T;                                   \
      F d = 0, alpha = 0, k = 1;                                                           \
      if (p >= 0.0) {                                                                      \
        d = (F)(d0 * pow(p, power));                                                       \
        alpha = (F)(alpha_max * (1.0 - p));                                                \
        k = (F)(1.0 + (kappa_max - 1.0) * p);                                              \
      }                                                                                    \
      F bb = 0, aa = 0;                                                                    \
      if (d > 0 || alpha > 0) {                                                            \
        const F dk = d / k;                                                                \
        const F s = dk + alpha;                                                            \
        bb = (F)exp(-(double)s * dt);                                                      \
        if (ABS(d) > EPS) aa = d / (k * (d + k * alpha)) * (bb - (F)1);                    \
      }                                                                                    \
      a[i] = aa;                                                                           \
      b[i] = bb;                                                                           \
      kappa[i] = k;                                                                        \
      if (dd) dd[i] = d;                                                                   \
      if (al) al[i] = alpha;                                                               \
    }                                                                                      \
  }
VIO_CPML_AXIS(vio_cpml_axis, double, DBL_EPSILON, fabs)
VIO_CPML_AXIS(vio_cpml_axis_f32, float, FLT_EPSILON, fabsf)
#undef VIO_CPML_AXIS

/* H-AFDA free surface (free_surface = 2): adjusted one-sided FD approximations after Kristek,
 * Moczo & Archuleta (2002, Stud. Geophys. Geod. 46), the scheme of the reference's draft
 * src/free-surface-kernels.jl:1-214 with src/common/Stencil.jl:52-105 (formulas #1, #2, #4),
 * carried to the rows an 8th-order centred stencil needs. The surface is row 4 (z = 0); row
 * 4+k holds sxx, szz, vx at z = k h and vz, sxz at z = (k+1/2) h; no value above z = 0 is
 * used. Each vertical derivative whose centred stencil would reach above the surface comes
 * from the first n rows below it (rows 4 .. 4+n-1 of the field):
 *   table rows  0.. 3  A_k  d sxz/dz at z = k h (k = 0..3): points 0 (sxz(0) = 0, its weight
 *                           dropped, formula #1/#4) and (m+1/2) h, m < n
 *   table rows  4.. 6  B_k  d szz/dz at z = (k+1/2) h: points m h, m < n (formula #2;
 *                           szz(0) = 0 is the stored surface value)
 *   table rows  7.. 9  C_k  d vz/dz at z = k h (k = 1..3): points (m+1/2) h, m < n
 *   table rows 10..12  D_k  d vx/dz at z = (k+1/2) h: points m h, m < n (formula #2)
 * The weights are those of the derivative of the Lagrange interpolant at x0,
 *   w_m = sum_{q != m} prod_{l != m, q} (x0 - x_l) / prod_{l != m} (x_m - x_l),
 * evaluated in integers of h/2 (every factor is at most 16, so numerator and denominator
 * are exact in int64 and in double) and rounded once by the final division: the correctly
 * rounded exact rational, then divided by h. Unused columns are 0. */
static const int vio_afda_default[13] = {3, 4, 6, 8, 3, 4, 6, 2, 4, 6, 3, 4, 6};

void vio_afda_default_points(int out[13]) {
  for (int r = 0; r < 13; ++r) out[r] = vio_afda_default[r];
}

int vio_afda_weights(double h, const int* counts, double out[13][8]) {
  const int* nc = counts ? counts : vio_afda_default;
  for (int r = 0; r < 13; ++r) {
    long long X[9], X0; /* positions in units of h/2 */
    int n = nc[r], skip = 0, np;
    if (n < (r < 4 ? 1 : 2) || n > 8) return -1;
    if (r < 4) {
      X[0] = 0;
      for (int m = 0; m < n; ++m) X[m + 1] = 2 * m + 1;
      np = n + 1;
      skip = 1;
      X0 = 2 * r;
    } else if (r < 10 && r >= 7) {
      for (int m = 0; m < n; ++m) X[m] = 2 * m + 1;
      np = n;
      X0 = 2 * (r - 7 + 1);
    } else {
      for (int m = 0; m < n; ++m) X[m] = 2 * m;
      np = n;
      X0 = 2 * (r < 7 ? r - 4 : r - 10) + 1;
    }
    for (int m = 0; m < 8; ++m) out[r][m] = 0.0;
    for (int m = skip; m < np; ++m) {
      long long num = 0, den = 1;
      for (int l = 0; l < np; ++l)
        if (l != m) den *= X[m] - X[l];
      for (int q = 0; q < np; ++q) {
        if (q == m) continue;
        long long t = 1;
        for (int l = 0; l < np; ++l)
          if (l != m && l != q) t *= X0 - X[l];
        num += t;
      }
      /* den carries one factor of h/2 more than num, hence 2 num / den per unit h */
      out[r][m - skip] = ((double)(2 * num) / (double)den) / h;
    }
  }
  return 0;
}

/* H-AFDA formula #3 (free_surface = 3): the Hermitian one-sided derivative of Stencil.jl:52-105,
 * 119-189 (_stencil_hermitian), d vz/dz at z = h (row 5) from the surface derivative f'(0)
 * and vz at the half points (m+1/2) h, m < n:
 *   f'(h) = alpha f'(0) + sum_m beta_m f((m+1/2) h),
 * exact for polynomials of degree <= n. With the nodes x_m = 2m+1 in units of h/2, w(x) =
 * prod (x - x_l) and l_m the Lagrange basis, every polynomial p of degree <= n is
 * sum p(x_m) l_m + c w, so p'(0) fixes c and
 *   alpha = w'(2) / w'(0),   beta_m = l_m'(2) - alpha l_m'(0)
 * (w'(0) != 0: all its terms have the sign (-1)^(n-1)). In integers, with W(x) = w'(x),
 * N_m(x) = sum_{q != m} prod_{l != m, q} (x - x_l) and D_m = prod_{l != m} (x_m - x_l):
 *   alpha = W(2) / W(0),   beta_m = 2 (N_m(2) W(0) - W(2) N_m(0)) / (D_m W(0)) per unit h,
 * numerators and denominators exact in int64 and in double for n <= 8 (below 2^53), each
 * rounded once by its division; out[0] = alpha, out[1 + m] = beta_m / h (0 for m >= n). */
int vio_afda_hermitian(double h, int n, double out[9]) {
  if (n < 2 || n > 8 || !(h > 0)) return -1;
  long long X[8];
  for (int m = 0; m < n; ++m) X[m] = 2 * m + 1;
  long long W0 = 0, W2 = 0, N0[8], N2[8];
  for (int k = 0; k < n; ++k) {
    long long t0 = 1, t2 = 1;
    for (int l = 0; l < n; ++l)
      if (l != k) { t0 *= -X[l]; t2 *= 2 - X[l]; }
    W0 += t0;
    W2 += t2;
    N0[k] = 0;
    N2[k] = 0;
    for (int q = 0; q < n; ++q) {
      if (q == k) continue;
      long long u0 = 1, u2 = 1;
      for (int l = 0; l < n; ++l)
        if (l != k && l != q) { u0 *= -X[l]; u2 *= 2 - X[l]; }
      N0[k] += u0;
      N2[k] += u2;
    }
  }
  for (int m = 0; m < 9; ++m) out[m] = 0.0;
  out[0] = (double)W2 / (double)W0;
  for (int k = 0; k < n; ++k) {
    long long D = 1;
    for (int l = 0; l < n; ++l)
      if (l != k) D *= X[k] - X[l];
    const long long num = 2 * (N2[k] * W0 - W2 * N0[k]), den = D * W0;
    out[1 + k] = ((double)num / (double)den) / h;
  }
  return 0;
}

#define REAL float
#define SUF _f32
#define VIO_CPML_AXIS_REAL vio_cpml_axis_f32
#include "vio_body.h"
#undef VIO_CPML_AXIS_REAL
#undef REAL
#undef SUF

#define REAL double
#define SUF _f64
#define VIO_CPML_AXIS_REAL vio_cpml_axis
#include "vio_body.h"
#undef VIO_CPML_AXIS_REAL
#undef REAL
#undef SUF
