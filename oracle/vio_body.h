/*
 * vio_body.h -- body of the CPU oracle, included by vio.c once per precision with
 * REAL (float | double) and SUF (_f32 | _f64) defined. TEST INFRASTRUCTURE ONLY.
 *
 * Discrete scheme (0-based indices, every field [nz][nx] with x fastest):
 *   positions (Virieux): sxx, szz, rho, lam, mu, tau at (i, j); vx at (i+1/2, j);
 *   vz at (i, j+1/2); sxz at (i+1/2, j+1/2).
 *   Interior I = [4, nx-4) x [4, nz-4) (visco-kernels.jl:95-96 pad = length(D) = 4);
 *   fields outside I are never updated (Dirichlet ring, BoundaryConditions.jl:8-11).
 *
 *   Operators, accumulated from l = 1 in REAL without contraction (diff-funcs.jl:3-32):
 *     DXP(u) = sum_l Dx[l] (u[i+l] - u[i-l+1])   DXM(u) = sum_l Dx[l] (u[i+l-1] - u[i-l])
 *     DZP, DZM the same along j.
 *
 *   Velocity half step n (visco-kernels.jl:3-20 with the Virieux operator choice):
 *     bx = 2/(rho[i]+rho[i+1]), bz = 2/(rho[j]+rho[j+1])          visco-kernels.jl:10-11
 *     fx = DXP(sxx) + DZM(sxz);  fz = DZP(szz) + DXM(sxz) (+ force_z samples)
 *     vx += dt*(bx*fx);  vz += dt*(bz*fz)
 *     force sources add their sample to fz before buoyancy   ViscoelasticProblem.jl:162-171
 *   Stress half step n (visco-kernels.jl:22-85):
 *     exx = DXM(vx), ezz = DZM(vz), gxz = DZP(vx) + DXP(vz)
 *     pi_r = lam + 2 mu; X1 = X*tau; X0 = X1 + X                  visco-kernels.jl:24-37
 *     mu_xz = 0.25*(((mu[i,j]+mu[i+1,j])+mu[i,j+1])+mu[i+1,j+1])  visco-kernels.jl:25
 *     dsxx = pi0 exx + lam0 ezz + pi1 SMxx + lam1 SMzz             visco-kernels.jl:42
 *     dszz = lam0 exx + pi0 ezz + lam1 SMxx + pi1 SMzz             free-surface-kernels.jl:197
 *     dsxz = mu0 gxz + mu1 SMxz
 *     s += dt*ds (old M), then M_n += dt*(a1_n e + a2_n M_n),
 *     a1_n = -1/(L taun_n), a2_n = -1/taun_n                        visco-kernels.jl:49-57
 *     explosive sources add their sample to dsxx and dszz.
 *   CPML (CPML.jl:91-189, ViscoelasticProblem.jl:208-216; a, b, kappa computed in REAL as
 *   the reference does, see vio.c), per derivative d in a strip:
 *     psi = b*psi + a*d;  d = d/kappa + psi   (profile index = array index of the cell)
 *   Free surface at row js = 4 (stress imaging, CPML.jl:372-380, generalised to 4 image
 *   rows for the 8th-order stencil): on row js, ezz is replaced by the value that makes
 *   dszz = 0, i.e. ezz = -((lam0 exx + lam1 SMxx) + pi1 SMzz)/pi0 (free-surface-kernels.jl:
 *   159-167 intent), and szz is set to 0 after the update; then
 *     szz[js-m] = -szz[js+m],  sxz[js-m] = -sxz[js+m-1],  m = 1..4, columns i in [4, nx-4).
 *   H-AFDA free surface (free_surface = 2; the one-sided weights are described in vio.c):
 *   the same surface row (ezz from dszz = 0, szz = 0), no image rows; instead, before the
 *   CPML correction, the vertical derivatives of the rows next to the surface come from the
 *   one-sided stencils on rows 4..11, each accumulated w0 f4 + w1 f5 + ... + w7 f11 in REAL:
 *     vx rows 4..7: DZM(sxz) -> A;  vz rows 4..6: DZP(szz) -> B;
 *     stress rows 5..7: DZM(vz) -> C;  sxz rows 4..6 (z = 1/2, 3/2, 5/2): DZP(vx) -> D.
 *   Receivers sample the stored value of their field at (ix, iz) after the stress half.
 *   VTI (VTIModel, Models.jl:27-33, with the TTI form ViscoelasticProblem.jl:21-35 at
 *   C15 = C35 = 0): lam holds C13, mu holds C55 (averaged to the sxz node like mu) and the
 *   diagonal moduli C11, C33 replace pi in the xx and zz rows:
 *     dsxx = C11_0 exx + C13_0 ezz + C11_1 SMxx + C13_1 SMzz
 *     dszz = C13_0 exx + C33_0 ezz + C13_1 SMxx + C33_1 SMzz,   X_1 = X tau, X_0 = X_1 + X
 *   and the free-surface rate is ezz = -((C13_0 exx + C13_1 SMxx) + C33_1 SMzz)/C33_0.
 *   With C11 = C33 = lam + 2 mu this is the isotropic update operation for operation.
 *
 * Defects of the draft resolved here: _ddz := ddzp (visco-kernels.jl:67 undefined); the
 * velocity kernel's operator choice is mirrored to match its own stress kernel
 * (visco-kernels.jl:13-16 vs 65-67); sxz uses the full engineering shear rate gxz
 * (the draft halves it, visco-kernels.jl:67 vs ViscoelasticProblem.jl:49); szz memory terms
 * use the symmetric form of free-surface-kernels.jl:197 (visco-kernels.jl:43 swaps them);
 * psi is updated (the loaded path never updates it, ViscoelasticProblem.jl:183-206).
 * Literals are cast to REAL, so an FP32 run computes in FP32 throughout.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define NAME(x) CAT(x, SUF)
#define SIM NAME(vio_sim)

#define VIO_MAXL 8

typedef struct {
  vio_desc d;
  int L, T;
  size_t ncell;
  REAL *vx, *vz, *sxx, *szz, *sxz;
  REAL *M; /* [3][L][nz][nx] */
  REAL *rho, *lam, *mu, *tau;
  REAL *c11, *c33; /* VTI only (NULL: isotropic, pi = lam + 2 mu) */
  REAL taun[VIO_MAXL], a1[VIO_MAXL], a2[VIO_MAXL];
  REAL Dx[4], Dz[4];
  REAL afda[13][8]; /* H-AFDA one-sided weights (free_surface = 2, 3) */
  REAL herm[9];     /* free_surface = 3: alpha, beta_m / h of formula #3 (d vz/dz on row 5) */
  REAL* es;         /* free_surface = 3: the surface's d vz/dz per column (fused path) */
  REAL *ax, *bx, *kx, *az, *bz, *kz;
  int xlo, xhi, zlo, zhi;
  REAL* psi[8];
  int nsrc, snt;
  int *six, *siz, *skind;
  REAL* sw;
  unsigned char* srcmark;
  int nrec;
  int *rix, *riz, *rcomp;
  long step;
  REAL* scr[7];
} SIM;

static REAL* NAME(vio_alloc)(size_t n) {
  REAL* p = (REAL*)calloc(n ? n : 1, sizeof(REAL));
  return p;
}

void* NAME(vio_create)(const vio_desc* d, const REAL* rho, const REAL* lam, const REAL* mu,
                       const REAL* tau, const REAL* taun) {
  if (!d || d->nx < 16 || d->nz < 16 || d->nrelax < 0 || d->nrelax > VIO_MAXL) return NULL;
  if (d->free_surface >= 2) {
    double w[13][8], hw[9];
    if (vio_afda_weights(d->dz, d->afda_points[0] ? d->afda_points : NULL, w)) return NULL;
    if (d->free_surface == 3 &&
        vio_afda_hermitian(d->dz, d->afda_hermitian_points ? d->afda_hermitian_points : VIO_HERMITIAN_DEFAULT, hw))
      return NULL;
  }
  {
    double c[4];
    const int ox = d->fd_order_x ? d->fd_order_x : 8, oz = d->fd_order_z ? d->fd_order_z : 8;
    if (vio_fd_weights(1.0, ox, c) || vio_fd_weights(1.0, oz, c)) return NULL;
    if (d->free_surface >= 2 && oz != 8) return NULL; /* the H-AFDA rows are 8th-order */
  }
  /* the surface row must not lie in a top CPML strip */
  if (d->free_surface && d->cpml_thickness > 0 && d->cpml_top) return NULL;
  SIM* s = (SIM*)calloc(1, sizeof(SIM));
  s->d = *d;
  s->L = d->nrelax;
  s->T = d->cpml_thickness > 0 ? d->cpml_thickness : 0;
  if (s->T && (2 * s->T > d->nx || 2 * s->T > d->nz)) { free(s); return NULL; }
  const int nx = d->nx, nz = d->nz;
  s->ncell = (size_t)nx * nz;
  s->vx = NAME(vio_alloc)(s->ncell); s->vz = NAME(vio_alloc)(s->ncell);
  s->sxx = NAME(vio_alloc)(s->ncell); s->szz = NAME(vio_alloc)(s->ncell);
  s->sxz = NAME(vio_alloc)(s->ncell);
  s->M = NAME(vio_alloc)((size_t)3 * s->L * s->ncell);
  s->rho = NAME(vio_alloc)(s->ncell); s->lam = NAME(vio_alloc)(s->ncell);
  s->mu = NAME(vio_alloc)(s->ncell); s->tau = NAME(vio_alloc)(s->ncell);
  memcpy(s->rho, rho, s->ncell * sizeof(REAL));
  memcpy(s->lam, lam, s->ncell * sizeof(REAL));
  memcpy(s->mu, mu, s->ncell * sizeof(REAL));
  if (s->L > 0) {
    memcpy(s->tau, tau, s->ncell * sizeof(REAL));
    for (int n = 0; n < s->L; ++n) {
      s->taun[n] = taun[n];
      REAL one = 1, Lr = (REAL)s->L;
      s->a1[n] = -one / (Lr * taun[n]);
      s->a2[n] = -one / taun[n];
    }
  }
  double c[4];
  vio_fd_weights(d->dx, d->fd_order_x ? d->fd_order_x : 8, c); /* validated above */
  for (int l = 0; l < 4; ++l) s->Dx[l] = (REAL)c[l];
  vio_fd_weights(d->dz, d->fd_order_z ? d->fd_order_z : 8, c);
  for (int l = 0; l < 4; ++l) s->Dz[l] = (REAL)c[l];
  if (d->free_surface >= 2) {
    double w[13][8];
    vio_afda_weights(d->dz, d->afda_points[0] ? d->afda_points : NULL, w); /* validated above */
    for (int r = 0; r < 13; ++r)
      for (int m = 0; m < 8; ++m) s->afda[r][m] = (REAL)w[r][m];
  }
  if (d->free_surface == 3) {
    double hw[9];
    vio_afda_hermitian(d->dz, d->afda_hermitian_points ? d->afda_hermitian_points : VIO_HERMITIAN_DEFAULT, hw);
    for (int m = 0; m < 9; ++m) s->herm[m] = (REAL)hw[m];
    s->es = NAME(vio_alloc)(nx);
  }

  s->ax = NAME(vio_alloc)(nx); s->bx = NAME(vio_alloc)(nx); s->kx = NAME(vio_alloc)(nx);
  s->az = NAME(vio_alloc)(nz); s->bz = NAME(vio_alloc)(nz); s->kz = NAME(vio_alloc)(nz);
  for (int i = 0; i < nx; ++i) s->kx[i] = 1;
  for (int j = 0; j < nz; ++j) s->kz[j] = 1;
  if (s->T) {
    s->xlo = d->cpml_left; s->xhi = d->cpml_right;
    s->zlo = d->cpml_top; s->zhi = d->cpml_bottom;
    /* coefficients in the working precision (CPML.jl:109-189) */
    VIO_CPML_AXIS_REAL(nx, s->T, d->cpml_d0, d->cpml_alpha_max, d->cpml_kappa_max,
                       d->cpml_power, d->dt, s->xlo, s->xhi, s->ax, s->bx, s->kx, NULL, NULL);
    VIO_CPML_AXIS_REAL(nz, s->T, d->cpml_d0, d->cpml_alpha_max, d->cpml_kappa_max,
                       d->cpml_power, d->dt, s->zlo, s->zhi, s->az, s->bz, s->kz, NULL, NULL);
    for (int q = 0; q < 4; ++q) s->psi[q] = NAME(vio_alloc)((size_t)nz * 2 * s->T);
    for (int q = 4; q < 8; ++q) s->psi[q] = NAME(vio_alloc)((size_t)2 * s->T * nx);
  }
  s->srcmark = (unsigned char*)calloc(s->ncell, 1);
  return s;
}

void NAME(vio_destroy)(void* h) {
  SIM* s = (SIM*)h;
  if (!s) return;
  free(s->vx); free(s->vz); free(s->sxx); free(s->szz); free(s->sxz); free(s->M);
  free(s->rho); free(s->lam); free(s->mu); free(s->tau); free(s->c11); free(s->c33);
  free(s->ax); free(s->bx); free(s->kx); free(s->az); free(s->bz); free(s->kz);
  for (int q = 0; q < 8; ++q) free(s->psi[q]);
  free(s->six); free(s->siz); free(s->skind); free(s->sw); free(s->srcmark);
  free(s->rix); free(s->riz); free(s->rcomp);
  for (int q = 0; q < 7; ++q) free(s->scr[q]);
  free(s->es);
  free(s);
}

long NAME(vio_step_index)(void* h) { return ((SIM*)h)->step; }

/* VTI instance: lam <- C13, mu <- C55, plus the C11 and C33 planes */
void* NAME(vio_create_vti)(const vio_desc* d, const REAL* rho, const REAL* c11, const REAL* c13,
                           const REAL* c33, const REAL* c55, const REAL* tau, const REAL* taun) {
  if (!c11 || !c33) return NULL;
  SIM* s = (SIM*)NAME(vio_create)(d, rho, c13, c55, tau, taun);
  if (!s) return NULL;
  s->c11 = NAME(vio_alloc)(s->ncell);
  s->c33 = NAME(vio_alloc)(s->ncell);
  memcpy(s->c11, c11, s->ncell * sizeof(REAL));
  memcpy(s->c33, c33, s->ncell * sizeof(REAL));
  return s;
}

static int NAME(in_interior)(const SIM* s, int i, int j) {
  return i >= 4 && i < s->d.nx - 4 && j >= 4 && j < s->d.nz - 4;
}

int NAME(vio_set_sources)(void* h, int n, const int* ix, const int* iz, const int* kind, int nt,
                          const REAL* w) {
  SIM* s = (SIM*)h;
  if (n < 0 || nt < 0) return -1;
  for (int q = 0; q < n; ++q)
    if (!NAME(in_interior)(s, ix[q], iz[q]) || kind[q] < 0 || kind[q] > 2) return -1;
  free(s->six); free(s->siz); free(s->skind); free(s->sw);
  memset(s->srcmark, 0, s->ncell);
  s->nsrc = n; s->snt = nt;
  s->six = (int*)malloc(sizeof(int) * (n ? n : 1));
  s->siz = (int*)malloc(sizeof(int) * (n ? n : 1));
  s->skind = (int*)malloc(sizeof(int) * (n ? n : 1));
  s->sw = NAME(vio_alloc)((size_t)n * nt);
  for (int q = 0; q < n; ++q) {
    s->six[q] = ix[q]; s->siz[q] = iz[q]; s->skind[q] = kind[q];
    s->srcmark[(size_t)iz[q] * s->d.nx + ix[q]] = 1;
  }
  memcpy(s->sw, w, sizeof(REAL) * (size_t)n * nt);
  return 0;
}

int NAME(vio_set_receivers)(void* h, int n, const int* ix, const int* iz, const int* comp) {
  SIM* s = (SIM*)h;
  if (n < 0) return -1;
  for (int q = 0; q < n; ++q)
    if (ix[q] < 0 || ix[q] >= s->d.nx || iz[q] < 0 || iz[q] >= s->d.nz || comp[q] < 0 ||
        comp[q] > 4)
      return -1;
  free(s->rix); free(s->riz); free(s->rcomp);
  s->nrec = n;
  s->rix = (int*)malloc(sizeof(int) * (n ? n : 1));
  s->riz = (int*)malloc(sizeof(int) * (n ? n : 1));
  s->rcomp = (int*)malloc(sizeof(int) * (n ? n : 1));
  memcpy(s->rix, ix, sizeof(int) * n);
  memcpy(s->riz, iz, sizeof(int) * n);
  memcpy(s->rcomp, comp, sizeof(int) * n);
  return 0;
}

static REAL* NAME(field_ptr)(SIM* s, int f, size_t* count) {
  const size_t nc = s->ncell;
  switch (f) {
    case VIO_VX: *count = nc; return s->vx;
    case VIO_VZ: *count = nc; return s->vz;
    case VIO_SXX: *count = nc; return s->sxx;
    case VIO_SZZ: *count = nc; return s->szz;
    case VIO_SXZ: *count = nc; return s->sxz;
    case VIO_MXX: case VIO_MZZ: case VIO_MXZ:
      if (!s->L) return NULL;
      *count = (size_t)s->L * nc;
      return s->M + (size_t)(f - VIO_MXX) * s->L * nc;
    default:
      if (f >= VIO_PSI_VX_X && f <= VIO_PSI_SZZ_Z) {
        if (!s->T) return NULL;
        *count = f < VIO_PSI_VX_Z ? (size_t)s->d.nz * 2 * s->T : (size_t)2 * s->T * s->d.nx;
        return s->psi[f - VIO_PSI_VX_X];
      }
      return NULL;
  }
}

int NAME(vio_get)(void* h, int field, REAL* out) {
  size_t n = 0;
  REAL* p = NAME(field_ptr)((SIM*)h, field, &n);
  if (!p) return -1;
  memcpy(out, p, n * sizeof(REAL));
  return 0;
}

int NAME(vio_set)(void* h, int field, const REAL* in) {
  size_t n = 0;
  REAL* p = NAME(field_ptr)((SIM*)h, field, &n);
  if (!p) return -1;
  memcpy(p, in, n * sizeof(REAL));
  return 0;
}

/* ---- operators -------------------------------------------------------------------- */
#define OP_P(u, c, st, D)                                                       \
  ({                                                                           \
    REAL v_ = (D)[0] * ((u)[(c) + (st)] - (u)[(c)]);                           \
    v_ = v_ + (D)[1] * ((u)[(c) + 2 * (st)] - (u)[(c) - (st)]);                \
    v_ = v_ + (D)[2] * ((u)[(c) + 3 * (st)] - (u)[(c) - 2 * (st)]);            \
    v_ = v_ + (D)[3] * ((u)[(c) + 4 * (st)] - (u)[(c) - 3 * (st)]);            \
    v_;                                                                        \
  })
#define OP_M(u, c, st, D)                                                       \
  ({                                                                           \
    REAL v_ = (D)[0] * ((u)[(c)] - (u)[(c) - (st)]);                           \
    v_ = v_ + (D)[1] * ((u)[(c) + (st)] - (u)[(c) - 2 * (st)]);                \
    v_ = v_ + (D)[2] * ((u)[(c) + 2 * (st)] - (u)[(c) - 3 * (st)]);            \
    v_ = v_ + (D)[3] * ((u)[(c) + 3 * (st)] - (u)[(c) - 4 * (st)]);            \
    v_;                                                                        \
  })

/* strip index of column i / row j, or -1 outside the enabled strips */
static inline int NAME(xk)(const SIM* s, int i) {
  if (!s->T) return -1;
  if (s->xlo && i < s->T) return i;
  if (s->xhi && i >= s->d.nx - s->T) return i - (s->d.nx - 2 * s->T);
  return -1;
}
static inline int NAME(zk)(const SIM* s, int j) {
  if (!s->T) return -1;
  if (s->zlo && j < s->T) return j;
  if (s->zhi && j >= s->d.nz - s->T) return j - (s->d.nz - 2 * s->T);
  return -1;
}

static inline REAL NAME(cpml)(REAL* psi, REAL a, REAL b, REAL k, REAL d) {
  REAL p = b * (*psi) + a * d;
  *psi = p;
  return d / k + p;
}

/* psi' = b psi + a d without storing it (the rates and the surface value of formula #3) */
static inline REAL NAME(cpml_rate)(const REAL* psi, REAL a, REAL b, REAL k, REAL d) {
  const REAL p = b * (*psi) + a * d;
  return d / k + p;
}

static inline REAL NAME(src_sample)(const SIM* s, int q, long n) {
  return n < s->snt ? s->sw[(size_t)q * s->snt + n] : (REAL)0;
}

/* one-sided d/dz at column i from rows 4..11 of u with the weights w */
static inline REAL NAME(afda_dz)(const REAL* u, int i, int nx, const REAL* w) {
  REAL v = w[0] * u[(size_t)4 * nx + i];
  for (int m = 1; m < 8; ++m) v = v + w[m] * u[(size_t)(4 + m) * nx + i];
  return v;
}

/* d vz/dz at the surface (row 4, column i) that keeps szz = 0 there: the free-surface row's
 * substitute ezz (dszz = 0), from the current state and without changing psi or M. With
 * L = 0 and isotropy it is -lam/(lam + 2 mu) d vx/dx, the w,z(0) of
 * free-surface-kernels.jl:160-172; the memory terms keep dszz = 0 when L > 0. */
static inline REAL NAME(surface_ezz)(const SIM* s, int i) {
  const int nx = s->d.nx, L = s->L;
  const size_t nc = s->ncell, c = (size_t)4 * nx + i;
  const REAL two = 2;
  REAL exx = OP_M(s->vx, c, 1, s->Dx);
  const int xk = NAME(xk)(s, i);
  if (xk >= 0)
    exx = NAME(cpml_rate)(&s->psi[0][(size_t)4 * 2 * s->T + xk], s->ax[i], s->bx[i], s->kx[i],
                          exx);
  const REAL lr = s->lam[c];
  const REAL pz = s->c11 ? s->c33[c] : lr + two * s->mu[c];
  if (L > 0) {
    const REAL t = s->tau[c];
    const REAL pz1 = pz * t, pz0 = pz1 + pz;
    const REAL l1 = lr * t, l0 = l1 + lr;
    const REAL* Mxx = s->M;
    const REAL* Mzz = s->M + (size_t)L * nc;
    REAL SMxx = Mxx[c], SMzz = Mzz[c];
    for (int k = 1; k < L; ++k) {
      SMxx = SMxx + Mxx[(size_t)k * nc + c];
      SMzz = SMzz + Mzz[(size_t)k * nc + c];
    }
    return -(((l0 * exx) + l1 * SMxx) + pz1 * SMzz) / pz0;
  }
  return -(lr * exx) / pz;
}

/* formula #3 on row 5: alpha vz,z(0) + sum_m beta_m vz(row 4 + m), in that order */
static inline REAL NAME(herm_dz)(const SIM* s, int i, REAL es) {
  const int nx = s->d.nx;
  REAL v = s->herm[0] * es;
  for (int m = 0; m < 8; ++m) v = v + s->herm[1 + m] * s->vz[(size_t)(4 + m) * nx + i];
  return v;
}

/* ---- fused path ------------------------------------------------------------------- */
static void NAME(vel_fused)(SIM* s, long n) {
  const int nx = s->d.nx, nz = s->d.nz, T = s->T;
  const int afda = s->d.free_surface >= 2;
  const REAL dt = (REAL)s->d.dt, two = 2;
  const REAL *Dx = s->Dx, *Dz = s->Dz;
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j) {
    const int zk = NAME(zk)(s, j);
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      REAL dsxx_x = OP_P(s->sxx, c, 1, Dx);
      REAL dsxz_z = afda && j < 8 ? NAME(afda_dz)(s->sxz, i, nx, s->afda[j - 4])
                                  : OP_M(s->sxz, c, (size_t)nx, Dz);
      REAL dsxz_x = OP_M(s->sxz, c, 1, Dx);
      REAL dszz_z = afda && j < 7 ? NAME(afda_dz)(s->szz, i, nx, s->afda[4 + j - 4])
                                  : OP_P(s->szz, c, (size_t)nx, Dz);
      const int xk = NAME(xk)(s, i);
      if (xk >= 0) {
        size_t p = (size_t)j * 2 * T + xk;
        dsxx_x = NAME(cpml)(&s->psi[2][p], s->ax[i], s->bx[i], s->kx[i], dsxx_x);
        dsxz_x = NAME(cpml)(&s->psi[3][p], s->ax[i], s->bx[i], s->kx[i], dsxz_x);
      }
      if (zk >= 0) {
        size_t p = (size_t)zk * nx + i;
        dsxz_z = NAME(cpml)(&s->psi[6][p], s->az[j], s->bz[j], s->kz[j], dsxz_z);
        dszz_z = NAME(cpml)(&s->psi[7][p], s->az[j], s->bz[j], s->kz[j], dszz_z);
      }
      REAL fx = dsxx_x + dsxz_z;
      REAL fz = dszz_z + dsxz_x;
      if (s->srcmark[c]) {
        for (int q = 0; q < s->nsrc; ++q) {
          if ((size_t)s->siz[q] * nx + s->six[q] != c) continue;
          if (s->skind[q] == VIO_SRC_FORCE_Z) fz = fz + NAME(src_sample)(s, q, n);
          else if (s->skind[q] == VIO_SRC_FORCE_X) fx = fx + NAME(src_sample)(s, q, n);
        }
      }
      const REAL bxv = two / (s->rho[c] + s->rho[c + 1]);
      const REAL bzv = two / (s->rho[c] + s->rho[c + nx]);
      s->vx[c] = s->vx[c] + dt * (bxv * fx);
      s->vz[c] = s->vz[c] + dt * (bzv * fz);
    }
  }
}

/* constitutive law + memory update for one cell; e* are CPML-corrected rates */
static inline void NAME(stress_cell)(SIM* s, size_t c, long n, REAL exx, REAL ezz, REAL gxz,
                                     int fs_row, REAL* ezz_used) {
  const int nx = s->d.nx, L = s->L;
  const REAL dt = (REAL)s->d.dt, two = 2, quarter = (REAL)0.25;
  const size_t nc = s->ncell;
  const REAL lr = s->lam[c], mr = s->mu[c];
  const REAL mxz = quarter * (((s->mu[c] + s->mu[c + 1]) + s->mu[c + nx]) + s->mu[c + nx + 1]);
  const REAL px = s->c11 ? s->c11[c] : lr + two * mr; /* C11 (pi when isotropic) */
  const REAL pz = s->c11 ? s->c33[c] : px;             /* C33 */
  REAL dsxx, dszz, dsxz;
  if (L > 0) {
    const REAL t = s->tau[c];
    const REAL px1 = px * t, px0 = px1 + px;
    const REAL pz1 = pz * t, pz0 = pz1 + pz;
    const REAL l1 = lr * t, l0 = l1 + lr;
    const REAL m1 = mxz * t, m0 = m1 + mxz;
    REAL* Mxx = s->M;
    REAL* Mzz = s->M + (size_t)L * nc;
    REAL* Mxz = s->M + (size_t)2 * L * nc;
    REAL SMxx = Mxx[c], SMzz = Mzz[c], SMxz = Mxz[c];
    for (int k = 1; k < L; ++k) {
      SMxx = SMxx + Mxx[(size_t)k * nc + c];
      SMzz = SMzz + Mzz[(size_t)k * nc + c];
      SMxz = SMxz + Mxz[(size_t)k * nc + c];
    }
    if (fs_row) ezz = -(((l0 * exx) + l1 * SMxx) + pz1 * SMzz) / pz0;
    dsxx = px0 * exx + l0 * ezz + px1 * SMxx + l1 * SMzz;
    dszz = l0 * exx + pz0 * ezz + l1 * SMxx + pz1 * SMzz;
    dsxz = m0 * gxz + m1 * SMxz;
  } else {
    if (fs_row) ezz = -(lr * exx) / pz;
    dsxx = px * exx + lr * ezz;
    dszz = lr * exx + pz * ezz;
    dsxz = mxz * gxz;
  }
  if (s->srcmark[c]) {
    for (int q = 0; q < s->nsrc; ++q) {
      if ((size_t)s->siz[q] * nx + s->six[q] != c || s->skind[q] != VIO_SRC_EXPLOSIVE) continue;
      const REAL w = NAME(src_sample)(s, q, n);
      dsxx = dsxx + w;
      dszz = dszz + w;
    }
  }
  s->sxx[c] = s->sxx[c] + dt * dsxx;
  s->szz[c] = s->szz[c] + dt * dszz;
  s->sxz[c] = s->sxz[c] + dt * dsxz;
  if (fs_row) s->szz[c] = 0;
  *ezz_used = ezz;
}

static inline void NAME(memory_cell)(SIM* s, size_t c, REAL exx, REAL ezz, REAL gxz) {
  const int L = s->L;
  const REAL dt = (REAL)s->d.dt;
  const size_t nc = s->ncell;
  REAL* Mxx = s->M;
  REAL* Mzz = s->M + (size_t)L * nc;
  REAL* Mxz = s->M + (size_t)2 * L * nc;
  for (int k = 0; k < L; ++k) {
    const size_t o = (size_t)k * nc + c;
    const REAL a1 = s->a1[k], a2 = s->a2[k];
    Mxx[o] = Mxx[o] + dt * (a1 * exx + a2 * Mxx[o]);
    Mzz[o] = Mzz[o] + dt * (a1 * ezz + a2 * Mzz[o]);
    Mxz[o] = Mxz[o] + dt * (a1 * gxz + a2 * Mxz[o]);
  }
}

static void NAME(images)(SIM* s) {
  const int nx = s->d.nx;
  for (int m = 1; m <= 4; ++m)
    for (int i = 4; i < nx - 4; ++i) {
      s->szz[(size_t)(4 - m) * nx + i] = -s->szz[(size_t)(4 + m) * nx + i];
      s->sxz[(size_t)(4 - m) * nx + i] = -s->sxz[(size_t)(4 + m - 1) * nx + i];
    }
}

static void NAME(stress_fused)(SIM* s, long n) {
  const int nx = s->d.nx, nz = s->d.nz, T = s->T;
  const REAL *Dx = s->Dx, *Dz = s->Dz;
  const int fs = s->d.free_surface;
  const int afda = fs >= 2, herm = fs == 3;
  /* formula #3 needs the surface value before row 4 updates psi and M */
  if (herm)
    for (int i = 4; i < nx - 4; ++i) s->es[i] = NAME(surface_ezz)(s, i);
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j) {
    const int zk = NAME(zk)(s, j);
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      REAL exx = OP_M(s->vx, c, 1, Dx);
      REAL ezz = herm && j == 5                 ? NAME(herm_dz)(s, i, s->es[i])
                 : afda && j >= 5 && j < 8 ? NAME(afda_dz)(s->vz, i, nx, s->afda[7 + j - 5])
                                           : OP_M(s->vz, c, (size_t)nx, Dz);
      REAL ezx = afda && j < 7 ? NAME(afda_dz)(s->vx, i, nx, s->afda[10 + j - 4])
                               : OP_P(s->vx, c, (size_t)nx, Dz);
      REAL exz = OP_P(s->vz, c, 1, Dx);
      const int xk = NAME(xk)(s, i);
      if (xk >= 0) {
        size_t p = (size_t)j * 2 * T + xk;
        exx = NAME(cpml)(&s->psi[0][p], s->ax[i], s->bx[i], s->kx[i], exx);
        exz = NAME(cpml)(&s->psi[1][p], s->ax[i], s->bx[i], s->kx[i], exz);
      }
      if (zk >= 0) {
        size_t p = (size_t)zk * nx + i;
        ezx = NAME(cpml)(&s->psi[4][p], s->az[j], s->bz[j], s->kz[j], ezx);
        ezz = NAME(cpml)(&s->psi[5][p], s->az[j], s->bz[j], s->kz[j], ezz);
      }
      const REAL gxz = ezx + exz;
      REAL ezz_used;
      NAME(stress_cell)(s, c, n, exx, ezz, gxz, fs && j == 4, &ezz_used);
      if (s->L) NAME(memory_cell)(s, c, exx, ezz_used, gxz);
    }
  }
  if (fs == 1) NAME(images)(s);
}

/* ---- reference-structured multi-pass path (ViscoelasticProblem.jl:62-174) ----------- */
static void NAME(ensure_scratch)(SIM* s) {
  for (int q = 0; q < 7; ++q)
    if (!s->scr[q]) s->scr[q] = NAME(vio_alloc)(s->ncell);
}

/* ddx!/ddz! analogue: one derivative into a scratch array over the padded interior */
static void NAME(deriv_pass)(SIM* s, REAL* out, const REAL* u, int plus, int along_z) {
  const int nx = s->d.nx, nz = s->d.nz;
  const REAL* D = along_z ? s->Dz : s->Dx;
  const size_t st = along_z ? (size_t)nx : 1;
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j)
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      out[c] = plus ? OP_P(u, c, st, D) : OP_M(u, c, st, D);
    }
}

/* H-AFDA rows: overwrite rows [j0, j1) of a d/dz scratch array with the one-sided stencils
 * whose table rows start at r0 */
static void NAME(afda_pass)(SIM* s, REAL* out, const REAL* u, int r0, int j0, int j1) {
  if (s->d.free_surface < 2) return;
  const int nx = s->d.nx;
  for (int j = j0; j < j1; ++j)
    for (int i = 4; i < nx - 4; ++i)
      out[(size_t)j * nx + i] = NAME(afda_dz)(u, i, nx, s->afda[r0 + j - j0]);
}

/* _cpml_scale_{x,z}! analogue with the psi recursion folded in */
static void NAME(cpml_pass)(SIM* s, REAL* du, REAL* psi, int along_z) {
  const int nx = s->d.nx, nz = s->d.nz, T = s->T;
  if (!T) return;
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j) {
    const int zk = NAME(zk)(s, j);
    if (along_z && zk < 0) continue;
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      if (along_z) {
        du[c] = NAME(cpml)(&psi[(size_t)zk * nx + i], s->az[j], s->bz[j], s->kz[j], du[c]);
      } else {
        const int xk = NAME(xk)(s, i);
        if (xk < 0) continue;
        du[c] = NAME(cpml)(&psi[(size_t)j * 2 * T + xk], s->ax[i], s->bx[i], s->kx[i], du[c]);
      }
    }
  }
}

static void NAME(vel_multipass)(SIM* s, long n) {
  const int nx = s->d.nx, nz = s->d.nz;
  const REAL dt = (REAL)s->d.dt, two = 2;
  NAME(ensure_scratch)(s);
  REAL *d1 = s->scr[0], *d2 = s->scr[1], *d3 = s->scr[2], *d4 = s->scr[3];
  NAME(deriv_pass)(s, d1, s->sxx, 1, 0); /* dx_sxx */
  NAME(deriv_pass)(s, d2, s->sxz, 0, 1); /* dz_sxz */
  NAME(deriv_pass)(s, d3, s->sxz, 0, 0); /* dx_szx */
  NAME(deriv_pass)(s, d4, s->szz, 1, 1); /* dz_szz */
  NAME(afda_pass)(s, d2, s->sxz, 0, 4, 8);
  NAME(afda_pass)(s, d4, s->szz, 4, 4, 7);
  NAME(cpml_pass)(s, d1, s->psi[2], 0);
  NAME(cpml_pass)(s, d3, s->psi[3], 0);
  NAME(cpml_pass)(s, d2, s->psi[6], 1);
  NAME(cpml_pass)(s, d4, s->psi[7], 1);
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j)
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      d1[c] = d1[c] + d2[c]; /* dvx = dx_sxx + dz_sxz */
      d4[c] = d4[c] + d3[c]; /* dvz = dz_szz + dx_szx */
    }
  for (int q = 0; q < s->nsrc; ++q) { /* source!(dvz, t) before buoyancy */
    const size_t c = (size_t)s->siz[q] * nx + s->six[q];
    if (s->skind[q] == VIO_SRC_FORCE_Z) d4[c] = d4[c] + NAME(src_sample)(s, q, n);
    else if (s->skind[q] == VIO_SRC_FORCE_X) d1[c] = d1[c] + NAME(src_sample)(s, q, n);
  }
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j)
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      d1[c] = (two / (s->rho[c] + s->rho[c + 1])) * d1[c];
      d4[c] = (two / (s->rho[c] + s->rho[c + nx])) * d4[c];
    }
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j)
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      s->vx[c] = s->vx[c] + dt * d1[c];
      s->vz[c] = s->vz[c] + dt * d4[c];
    }
}

static void NAME(stress_multipass)(SIM* s, long n) {
  const int nx = s->d.nx, nz = s->d.nz, L = s->L;
  const REAL dt = (REAL)s->d.dt, two = 2, quarter = (REAL)0.25;
  const size_t nc = s->ncell;
  const int fs = s->d.free_surface;
  NAME(ensure_scratch)(s);
  REAL *e1 = s->scr[0], *e2 = s->scr[1], *e3 = s->scr[2], *e4 = s->scr[3];
  REAL *ds1 = s->scr[4], *ds2 = s->scr[5], *ds3 = s->scr[6];
  NAME(deriv_pass)(s, e1, s->vx, 0, 0); /* dxvx */
  NAME(deriv_pass)(s, e2, s->vz, 0, 1); /* dzvz */
  NAME(deriv_pass)(s, e3, s->vx, 1, 1); /* dzvx */
  NAME(deriv_pass)(s, e4, s->vz, 1, 0); /* dxvz */
  NAME(afda_pass)(s, e2, s->vz, 7, 5, 8);
  NAME(afda_pass)(s, e3, s->vx, 10, 4, 7);
  if (fs == 3) /* formula #3 on row 5, before psi and M change */
    for (int i = 4; i < nx - 4; ++i)
      e2[(size_t)5 * nx + i] = NAME(herm_dz)(s, i, NAME(surface_ezz)(s, i));
  NAME(cpml_pass)(s, e1, s->psi[0], 0);
  NAME(cpml_pass)(s, e4, s->psi[1], 0);
  NAME(cpml_pass)(s, e3, s->psi[4], 1);
  NAME(cpml_pass)(s, e2, s->psi[5], 1);
  REAL* Mxx = s->M;
  REAL* Mzz = s->M + (size_t)L * nc;
  REAL* Mxz = s->M + (size_t)2 * L * nc;
  /* _update_ds! analogue (ViscoelasticProblem.jl:38-50) extended with the SLS terms */
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j)
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      const REAL exx = e1[c];
      REAL ezz = e2[c];
      const REAL gxz = e3[c] + e4[c];
      e3[c] = gxz;
      const REAL lr = s->lam[c], mr = s->mu[c];
      const REAL mxz = quarter * (((s->mu[c] + s->mu[c + 1]) + s->mu[c + nx]) + s->mu[c + nx + 1]);
      const REAL px = s->c11 ? s->c11[c] : lr + two * mr;
      const REAL pz = s->c11 ? s->c33[c] : px;
      const int fs_row = fs && j == 4;
      if (L > 0) {
        const REAL t = s->tau[c];
        const REAL px1 = px * t, px0 = px1 + px, pz1 = pz * t, pz0 = pz1 + pz;
        const REAL l1 = lr * t, l0 = l1 + lr, m1 = mxz * t, m0 = m1 + mxz;
        REAL SMxx = Mxx[c], SMzz = Mzz[c], SMxz = Mxz[c];
        for (int k = 1; k < L; ++k) {
          SMxx = SMxx + Mxx[(size_t)k * nc + c];
          SMzz = SMzz + Mzz[(size_t)k * nc + c];
          SMxz = SMxz + Mxz[(size_t)k * nc + c];
        }
        if (fs_row) ezz = -(((l0 * exx) + l1 * SMxx) + pz1 * SMzz) / pz0;
        ds1[c] = px0 * exx + l0 * ezz + px1 * SMxx + l1 * SMzz;
        ds2[c] = l0 * exx + pz0 * ezz + l1 * SMxx + pz1 * SMzz;
        ds3[c] = m0 * gxz + m1 * SMxz;
      } else {
        if (fs_row) ezz = -(lr * exx) / pz;
        ds1[c] = px * exx + lr * ezz;
        ds2[c] = lr * exx + pz * ezz;
        ds3[c] = mxz * gxz;
      }
      e2[c] = ezz;
    }
  for (int q = 0; q < s->nsrc; ++q) {
    if (s->skind[q] != VIO_SRC_EXPLOSIVE) continue;
    const size_t c = (size_t)s->siz[q] * nx + s->six[q];
    const REAL w = NAME(src_sample)(s, q, n);
    ds1[c] = ds1[c] + w;
    ds2[c] = ds2[c] + w;
  }
#pragma omp parallel for schedule(static)
  for (int j = 4; j < nz - 4; ++j)
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      s->sxx[c] = s->sxx[c] + dt * ds1[c];
      s->szz[c] = s->szz[c] + dt * ds2[c];
      s->sxz[c] = s->sxz[c] + dt * ds3[c];
      if (fs && j == 4) s->szz[c] = 0;
    }
  if (L > 0) {
#pragma omp parallel for schedule(static)
    for (int j = 4; j < nz - 4; ++j)
      for (int i = 4; i < nx - 4; ++i) {
        const size_t c = (size_t)j * nx + i;
        NAME(memory_cell)(s, c, e1[c], e2[c], e3[c]);
      }
  }
  if (fs == 1) NAME(images)(s);
}

static void NAME(sample)(SIM* s, REAL* out) {
  const int nx = s->d.nx;
  for (int r = 0; r < s->nrec; ++r) {
    const size_t c = (size_t)s->riz[r] * nx + s->rix[r];
    const REAL* f = s->rcomp[r] == 0   ? s->vx
                    : s->rcomp[r] == 1 ? s->vz
                    : s->rcomp[r] == 2 ? s->sxx
                    : s->rcomp[r] == 3 ? s->szz
                                       : s->sxz;
    out[r] = f[c];
  }
}

int NAME(vio_run)(void* h, int nsteps, int mode, int nthreads, REAL* traces) {
  SIM* s = (SIM*)h;
  if (!s || nsteps < 0) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int k = 0; k < nsteps; ++k) {
    const long n = s->step;
    if (mode == VIO_MULTIPASS) {
      NAME(vel_multipass)(s, n);
      NAME(stress_multipass)(s, n);
    } else {
      NAME(vel_fused)(s, n);
      NAME(stress_fused)(s, n);
    }
    if (traces && s->nrec) NAME(sample)(s, traces + (size_t)k * s->nrec);
    s->step = n + 1;
  }
  return 0;
}

/* ---- right-hand side of the step (stress_eq! / velocity_eq!, ViscoelasticProblem.jl:62-174) --
 * The rates the next half steps apply, from the current state, written without changing it:
 *   dv[2][nz][nx] = (vx, vz) rates of the velocity half step from the stored stresses (image
 *     rows as stored), force sources from src_values[q] (the sample of source q; NULL = 0);
 *   ds[3][nz][nx] = (sxx, szz, sxz) rates of the stress half step from the stored velocities
 *     and memory variables, explosive sources from src_values; on the free-surface row the
 *     ezz that makes dszz = 0 is used and dszz is 0; with stress imaging the image rows carry
 *     the negated rates of their mirror rows;
 *   dm[3][L][nz][nx] (may be NULL) = a1_n e + a2_n M_n per mechanism.
 * In CPML strips each derivative uses psi' = b psi + a d (psi itself is not changed), as the
 * step does. Cells outside the interior get 0. So the step's update is x + dt * rate in REAL,
 * operation for operation. */
int NAME(vio_rates)(void* h, const REAL* src_values, REAL* dv, REAL* ds, REAL* dm) {
  SIM* s = (SIM*)h;
  if (!s || !dv || !ds) return -1;
  const int nx = s->d.nx, nz = s->d.nz, T = s->T, L = s->L;
  const size_t nc = s->ncell;
  const int fs = s->d.free_surface, afda = fs >= 2, herm = fs == 3;
  const REAL two = 2, quarter = (REAL)0.25;
  const REAL *Dx = s->Dx, *Dz = s->Dz;
  memset(dv, 0, 2 * nc * sizeof(REAL));
  memset(ds, 0, 3 * nc * sizeof(REAL));
  if (dm && L) memset(dm, 0, (size_t)3 * L * nc * sizeof(REAL));
  REAL *dvx = dv, *dvz = dv + nc, *dsxx = ds, *dszz = ds + nc, *dsxz = ds + 2 * nc;
  for (int j = 4; j < nz - 4; ++j) {
    const int zk = NAME(zk)(s, j);
    for (int i = 4; i < nx - 4; ++i) {
      const size_t c = (size_t)j * nx + i;
      const int xk = NAME(xk)(s, i);
      /* velocity rates */
      REAL dsxx_x = OP_P(s->sxx, c, 1, Dx);
      REAL dsxz_z = afda && j < 8 ? NAME(afda_dz)(s->sxz, i, nx, s->afda[j - 4])
                                  : OP_M(s->sxz, c, (size_t)nx, Dz);
      REAL dsxz_x = OP_M(s->sxz, c, 1, Dx);
      REAL dszz_z = afda && j < 7 ? NAME(afda_dz)(s->szz, i, nx, s->afda[4 + j - 4])
                                  : OP_P(s->szz, c, (size_t)nx, Dz);
      if (xk >= 0) {
        const size_t p = (size_t)j * 2 * T + xk;
        dsxx_x = NAME(cpml_rate)(&s->psi[2][p], s->ax[i], s->bx[i], s->kx[i], dsxx_x);
        dsxz_x = NAME(cpml_rate)(&s->psi[3][p], s->ax[i], s->bx[i], s->kx[i], dsxz_x);
      }
      if (zk >= 0) {
        const size_t p = (size_t)zk * nx + i;
        dsxz_z = NAME(cpml_rate)(&s->psi[6][p], s->az[j], s->bz[j], s->kz[j], dsxz_z);
        dszz_z = NAME(cpml_rate)(&s->psi[7][p], s->az[j], s->bz[j], s->kz[j], dszz_z);
      }
      REAL fx = dsxx_x + dsxz_z;
      REAL fz = dszz_z + dsxz_x;
      int has_x = 0; /* an explosive source on this cell */
      if (s->srcmark[c]) {
        for (int q = 0; q < s->nsrc; ++q) {
          if ((size_t)s->siz[q] * nx + s->six[q] != c) continue;
          const REAL w = src_values ? src_values[q] : (REAL)0;
          if (s->skind[q] == VIO_SRC_FORCE_Z) fz = fz + w;
          else if (s->skind[q] == VIO_SRC_FORCE_X) fx = fx + w;
          else has_x = 1;
        }
      }
      dvx[c] = (two / (s->rho[c] + s->rho[c + 1])) * fx;
      dvz[c] = (two / (s->rho[c] + s->rho[c + nx])) * fz;
      /* stress rates */
      REAL exx = OP_M(s->vx, c, 1, Dx);
      REAL ezz = herm && j == 5                 ? NAME(herm_dz)(s, i, NAME(surface_ezz)(s, i))
                 : afda && j >= 5 && j < 8 ? NAME(afda_dz)(s->vz, i, nx, s->afda[7 + j - 5])
                                           : OP_M(s->vz, c, (size_t)nx, Dz);
      REAL ezx = afda && j < 7 ? NAME(afda_dz)(s->vx, i, nx, s->afda[10 + j - 4])
                               : OP_P(s->vx, c, (size_t)nx, Dz);
      REAL exz = OP_P(s->vz, c, 1, Dx);
      if (xk >= 0) {
        const size_t p = (size_t)j * 2 * T + xk;
        exx = NAME(cpml_rate)(&s->psi[0][p], s->ax[i], s->bx[i], s->kx[i], exx);
        exz = NAME(cpml_rate)(&s->psi[1][p], s->ax[i], s->bx[i], s->kx[i], exz);
      }
      if (zk >= 0) {
        const size_t p = (size_t)zk * nx + i;
        ezx = NAME(cpml_rate)(&s->psi[4][p], s->az[j], s->bz[j], s->kz[j], ezx);
        ezz = NAME(cpml_rate)(&s->psi[5][p], s->az[j], s->bz[j], s->kz[j], ezz);
      }
      const REAL gxz = ezx + exz;
      const REAL lr = s->lam[c], mr = s->mu[c];
      const REAL mxz = quarter * (((s->mu[c] + s->mu[c + 1]) + s->mu[c + nx]) + s->mu[c + nx + 1]);
      const REAL px = s->c11 ? s->c11[c] : lr + two * mr;
      const REAL pz = s->c11 ? s->c33[c] : px;
      const int fs_row = fs && j == 4;
      REAL r_xx, r_zz, r_xz;
      if (L > 0) {
        const REAL t = s->tau[c];
        const REAL px1 = px * t, px0 = px1 + px, pz1 = pz * t, pz0 = pz1 + pz;
        const REAL l1 = lr * t, l0 = l1 + lr, m1 = mxz * t, m0 = m1 + mxz;
        const REAL* Mxx = s->M;
        const REAL* Mzz = s->M + (size_t)L * nc;
        const REAL* Mxz = s->M + (size_t)2 * L * nc;
        REAL SMxx = Mxx[c], SMzz = Mzz[c], SMxz = Mxz[c];
        for (int k = 1; k < L; ++k) {
          SMxx = SMxx + Mxx[(size_t)k * nc + c];
          SMzz = SMzz + Mzz[(size_t)k * nc + c];
          SMxz = SMxz + Mxz[(size_t)k * nc + c];
        }
        if (fs_row) ezz = -(((l0 * exx) + l1 * SMxx) + pz1 * SMzz) / pz0;
        r_xx = px0 * exx + l0 * ezz + px1 * SMxx + l1 * SMzz;
        r_zz = l0 * exx + pz0 * ezz + l1 * SMxx + pz1 * SMzz;
        r_xz = m0 * gxz + m1 * SMxz;
        if (dm) {
          REAL* o = dm;
          for (int k = 0; k < L; ++k) {
            const size_t q = (size_t)k * nc + c;
            o[q] = s->a1[k] * exx + s->a2[k] * Mxx[q];
            o[(size_t)L * nc + q] = s->a1[k] * ezz + s->a2[k] * Mzz[q];
            o[(size_t)2 * L * nc + q] = s->a1[k] * gxz + s->a2[k] * Mxz[q];
          }
        }
      } else {
        if (fs_row) ezz = -(lr * exx) / pz;
        r_xx = px * exx + lr * ezz;
        r_zz = lr * exx + pz * ezz;
        r_xz = mxz * gxz;
      }
      if (has_x) {
        /* explosive sources, each added in turn as the step does */
        for (int q = 0; q < s->nsrc; ++q) {
          if ((size_t)s->siz[q] * nx + s->six[q] != c || s->skind[q] != VIO_SRC_EXPLOSIVE) continue;
          const REAL w = src_values ? src_values[q] : (REAL)0;
          r_xx = r_xx + w;
          r_zz = r_zz + w;
        }
      }
      dsxx[c] = r_xx;
      dszz[c] = fs_row ? (REAL)0 : r_zz;
      dsxz[c] = r_xz;
    }
  }
  if (fs == 1)
    for (int m = 1; m <= 4; ++m)
      for (int i = 4; i < nx - 4; ++i) {
        dszz[(size_t)(4 - m) * nx + i] = -dszz[(size_t)(4 + m) * nx + i];
        dsxz[(size_t)(4 - m) * nx + i] = -dsxz[(size_t)(4 + m - 1) * nx + i];
      }
  return 0;
}

#undef SIM
#undef NAME
#undef CAT
#undef CAT_
#undef OP_P
#undef OP_M
#undef VIO_MAXL
