/*
 * vio.h -- CPU oracle for the 2D staggered-grid viscoelastic velocity-stress step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product path
 * (libvi_mi355x.so) is compared against, and the CPU baseline that bench.py times beside
 * it. Nothing in the product links, loads or calls it.
 *
 * It restates, in plain C (one body compiled for float and for double), the scheme of
 * ViscoelasticInversion's draft fused kernels and loaded solver:
 *   - staggered 8th-order operators ddxp/ddxm/ddzp/ddzm      src/diff-funcs.jl:3-32
 *   - fused velocity update (buoyancy averaging)              src/visco-kernels.jl:3-20
 *   - SLS constitutive update + memory variables              src/visco-kernels.jl:22-85
 *   - CPML profiles and recursive-convolution coefficients    src/sciml_solver/CPML.jl:91-189
 *   - derivative stretching du/kappa + psi                    src/sciml_solver/ViscoelasticProblem.jl:208-216
 *   - Ricker force source added to dvz before buoyancy        src/sciml_solver/ViscoelasticProblem.jl:162-171
 *   - stress imaging free surface                             src/sciml_solver/CPML.jl:372-380
 *   - H-AFDA free surface (one-sided stencils, Kristek 2002)  src/free-surface-kernels.jl:1-214,
 *                                                             src/common/Stencil.jl:52-105
 *   - padded (Dirichlet) interior ranges                      src/common/Stencil.jl:432-447, visco-kernels.jl:95-96
 * with the draft defects resolved as listed in the oracle's header comment of vio_body.h.
 *
 * Two independent code paths are provided: VIO_FUSED (one pass per half step, like the
 * fused draft kernels) and VIO_MULTIPASS (derivatives into scratch arrays, then boundary,
 * CPML, constitutive and buoyancy passes, like ViscoelasticProblem.jl:62-174). Both use
 * the same floating-point operation order and must agree bit for bit.
 *
 * Layout: every field is a [nz][nx] array with x fastest (Julia's column-major (Nx,Nz)).
 */
#ifndef VIO_H
#define VIO_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VIO_VX = 0, VIO_VZ = 1, VIO_SXX = 2, VIO_SZZ = 3, VIO_SXZ = 4,
  VIO_MXX = 5, VIO_MZZ = 6, VIO_MXZ = 7,          /* L planes each, [L][nz][nx] */
  VIO_PSI_VX_X = 8, VIO_PSI_VZ_X = 9, VIO_PSI_SXX_X = 10, VIO_PSI_SXZ_X = 11, /* [nz][2T] */
  VIO_PSI_VX_Z = 12, VIO_PSI_VZ_Z = 13, VIO_PSI_SXZ_Z = 14, VIO_PSI_SZZ_Z = 15, /* [2T][nx] */
  VIO_NFIELDS = 16
};

enum { VIO_SRC_FORCE_Z = 0, VIO_SRC_EXPLOSIVE = 1, VIO_SRC_FORCE_X = 2 };
enum { VIO_FUSED = 0, VIO_MULTIPASS = 1 };

typedef struct {
  int nx, nz;          /* cells */
  int nrelax;          /* L: number of SLS mechanisms, 0 = elastic */
  double dx, dz, dt;
  int free_surface;    /* traction-free surface at row 4: 0 none, 1 stress images in rows
                          0..3, 2 H-AFDA one-sided stencils (nothing above the surface),
                          3 H-AFDA with the Hermitian formula #3 for d vz/dz on row 5 */
  int cpml_thickness;  /* 0: no CPML */
  double cpml_d0, cpml_alpha_max, cpml_kappa_max, cpml_power;
  int cpml_left, cpml_right, cpml_top, cpml_bottom;
  int afda_points[13]; /* H-AFDA point counts per table row (see vio.c); all 0 = default */
  int afda_hermitian_points; /* free_surface = 3: half points of formula #3; 0 = default */
  int fd_order_x, fd_order_z; /* staggered operator orders 2, 4, 6, 8; 0 = 8 */
} vio_desc;

/* Exact staggered 8th-order weights (1225/1024, -245/3072, 49/5120, -5/7168) / h. */
void vio_fd_coeffs(double h, double out[4]);
/* Staggered weights / h of order 2, 4, 6 or 8 (entries beyond order/2 are 0); -1 if invalid. */
int vio_fd_weights(double h, int order, double out[4]);

/* H-AFDA one-sided d/dz weights / h (free_surface = 2): table rows A0..A3, B0..B2, C1..C3,
 * D0..D2 over field rows 4 .. 11 (see vio.c). counts NULL = the default point counts.
 * Returns -1 for a count out of range. */
int vio_afda_weights(double h, const int* counts, double out[13][8]);
void vio_afda_default_points(int out[13]);

/* H-AFDA formula #3 (free_surface = 3): out[0] = alpha (multiplies d vz/dz at the surface),
 * out[1 + m] = beta_m / h (multiplies vz at (m+1/2) h, m < n); d vz/dz at z = h. Returns -1
 * for n outside [2, 8]. */
int vio_afda_hermitian(double h, int n, double out[9]);
/* default half-point count of formula #3: the largest that stays bounded in the measured
 * heterogeneous media (6, 7 and the reference's 8 grow without bound) */
#define VIO_HERMITIAN_DEFAULT 5

/* CPML a, b, kappa (and the profiles d, alpha; NULL to skip) along one axis of n cells
 * (CPML.jl:109-189) in double and in float, each computed in its own precision as the
 * reference computes them for Float64 and Float32 arrays; lo/hi = strip enabled. */
void vio_cpml_axis(int n, int thickness, double d0, double alpha_max, double kappa_max,
                   double power, double dt, int lo, int hi,
                   double* a, double* b, double* kappa, double* d, double* alpha);
void vio_cpml_axis_f32(int n, int thickness, double d0, double alpha_max, double kappa_max,
                       double power, double dt, int lo, int hi,
                       float* a, float* b, float* kappa, float* d, float* alpha);

/* float instance */
void* vio_create_f32(const vio_desc* d, const float* rho, const float* lam, const float* mu,
                     const float* tau, const float* taun);
void* vio_create_vti_f32(const vio_desc* d, const float* rho, const float* c11, const float* c13,
                         const float* c33, const float* c55, const float* tau, const float* taun);
int vio_set_sources_f32(void* h, int n, const int* ix, const int* iz, const int* kind,
                        int nt, const float* wavelets);
int vio_set_receivers_f32(void* h, int n, const int* ix, const int* iz, const int* comp);
int vio_run_f32(void* h, int nsteps, int mode, int nthreads, float* traces);
int vio_get_f32(void* h, int field, float* out);
int vio_set_f32(void* h, int field, const float* in);
long vio_step_index_f32(void* h);
/* rates of the next half steps at the current state (see vio_body.h): dv[2][nz][nx],
 * ds[3][nz][nx], dm[3][L][nz][nx] (may be NULL); src_values[nsrc] (may be NULL) */
int vio_rates_f32(void* h, const float* src_values, float* dv, float* ds, float* dm);
void vio_destroy_f32(void* h);

/* double instance */
void* vio_create_f64(const vio_desc* d, const double* rho, const double* lam, const double* mu,
                     const double* tau, const double* taun);
void* vio_create_vti_f64(const vio_desc* d, const double* rho, const double* c11,
                         const double* c13, const double* c33, const double* c55,
                         const double* tau, const double* taun);
int vio_set_sources_f64(void* h, int n, const int* ix, const int* iz, const int* kind,
                        int nt, const double* wavelets);
int vio_set_receivers_f64(void* h, int n, const int* ix, const int* iz, const int* comp);
int vio_run_f64(void* h, int nsteps, int mode, int nthreads, double* traces);
int vio_get_f64(void* h, int field, double* out);
int vio_set_f64(void* h, int field, const double* in);
long vio_step_index_f64(void* h);
int vio_rates_f64(void* h, const double* src_values, double* dv, double* ds, double* dm);
void vio_destroy_f64(void* h);

#ifdef __cplusplus
}
#endif
#endif
