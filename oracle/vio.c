/*
 * vio.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see vio.h). Compiled with
 * -ffp-contract=off so that every multiply and add rounds separately, as Julia does
 * without @fastmath/muladd (the reference's arithmetic), and so that the HIP kernels,
 * built with the same flag, can be compared with it bit for bit.
 */
#include "vio.h"

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

/* CPMLCoefficients (src/sciml_solver/CPML.jl:109-189) along one axis, in double:
 *   lo strip i in [0,T):       p = (T - i - 0.5)/T
 *   hi strip i in [n-T, n):    p = (i - (n-T) + 0.5)/T
 *   d = d0 p^power, alpha = alpha_max (1-p), kappa = 1 + (kappa_max-1) p
 *   b = exp(-(d/kappa + alpha) dt)                 where d > 0 or alpha > 0, else 0
 *   a = d/(kappa (d + kappa alpha)) (b - 1)        where |d| > 0, else 0            */
void vio_cpml_axis(int n, int T, double d0, double alpha_max, double kappa_max, double power,
                   double dt, int lo, int hi, double* a, double* b, double* kappa, double* dd,
                   double* al) {
  for (int i = 0; i < n; ++i) {
    double p = -1.0;
    if (lo && i < T) p = (T - i - 0.5) / T;
    if (hi && i >= n - T) p = (i - (n - T) + 0.5) / T;
    double d = 0.0, alpha = 0.0, k = 1.0;
    if (p >= 0.0) {
      d = d0 * pow(p, power);
      alpha = alpha_max * (1.0 - p);
      k = 1.0 + (kappa_max - 1.0) * p;
    }
    double bb = 0.0, aa = 0.0;
    if (d > 0.0 || alpha > 0.0) {
      bb = exp(-(d / k + alpha) * dt);
      if (fabs(d) > 0.0) aa = d / (k * (d + k * alpha)) * (bb - 1.0);
    }
    a[i] = aa;
    b[i] = bb;
    kappa[i] = k;
    if (dd) dd[i] = d;
    if (al) al[i] = alpha;
  }
}

#define REAL float
#define SUF _f32
#include "vio_body.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF _f64
#include "vio_body.h"
#undef REAL
#undef SUF
