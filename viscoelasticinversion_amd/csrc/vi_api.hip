// vi_api.hip -- host side of the C ABI declared in include/vi_mi355x.h.
//
// A handle (vi_sim_s) owns the device arrays of one simulation in the padded layout the
// kernels expect, one HIP stream, and the per-tile source/receiver tables. The precision
// is a run-time property of the handle; the implementation is the template Sim<R>.
#include "vi_mi355x.h"
#include "vi_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define VI_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(e_ == hipErrorOutOfMemory ? VI_ENOMEM : VI_EDEVICE,              \
                  std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// CPMLCoefficients along one axis (CPML.jl:109-189), in double.
void cpml_axis(int n, int T, double d0, double alpha_max, double kappa_max, double power,
               double dt, int lo, int hi, double* a, double* b, double* kappa) {
  for (int i = 0; i < n; ++i) {
    double p = -1.0;
    if (lo && i < T) p = (T - i - 0.5) / T;             // CPML.jl:124
    if (hi && i >= n - T) p = (i - (n - T) + 0.5) / T;  // CPML.jl:134
    double d = 0.0, alpha = 0.0, k = 1.0;
    if (p >= 0.0) {
      d = d0 * std::pow(p, power);          // CPML.jl:91-93
      alpha = alpha_max * (1.0 - p);        // CPML.jl:95-98
      k = 1.0 + (kappa_max - 1.0) * p;      // CPML.jl:100-102
    }
    double bb = 0.0, aa = 0.0;
    if (d > 0.0 || alpha > 0.0) {
      bb = std::exp(-(d / k + alpha) * dt);                           // CPML.jl:174
      if (std::fabs(d) > 0.0) aa = d / (k * (d + k * alpha)) * (bb - 1.0);  // CPML.jl:176
    }
    a[i] = aa;
    b[i] = bb;
    kappa[i] = k;
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t nbytes, hipStream_t s) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    bytes = nbytes;
    if (!nbytes) return 0;
    VI_HIP(hipMalloc(&p, nbytes));
    VI_HIP(hipMemsetAsync(p, 0, nbytes, s));
    return 0;
  }
};

struct SimBase {
  virtual ~SimBase() = default;
};

}  // namespace

struct vi_sim_s {
  vi_grid_desc g{};
  int L = 0;
  int ldx = 0, nzp = 0;
  size_t plane = 0;
  int ntx = 0, ntz = 0, ntiles = 0;
  int rz = 8;  // rows per wavefront of the current kernel variant
  int device = 0;
  hipStream_t stream = nullptr;
  size_t esz = 4;
  DevBuf state[5];  // vx vz sxx szz sxz
  DevBuf M;
  DevBuf model[4];  // rho lam mu tau
  DevBuf psi[8];
  DevBuf coef;      // ax bx kx (nx each), az bz kz (nz each)
  DevBuf src_tab, src_w, rec_tab, traces;
  bool model_set = false;
  double taun[vi::MAXL] = {};
  bool relax_set = false;
  bool cpml_on = false;
  vi_cpml_desc cpml{};
  int fs = 0;
  int nsrc = 0, snt = 0;
  int nrec = 0, rec_nt = 0;
  std::vector<int> src_ix, src_iz, src_kind, rec_ix, rec_iz, rec_comp;  // host copies
  long step = 0;
  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev;  // pairs
  std::vector<int> ev_kind;
  size_t ev_used = 0;
  double prof_ms[2] = {0, 0};
  long prof_n[2] = {0, 0};
};

namespace {

int check_handle(vi_handle h) {
  if (!h) return fail(VI_EINVAL, "null handle");
  return 0;
}

// Upload a host [nz][nx] array into a padded device plane.
int upload_plane(vi_sim_s* h, void* dst, const void* src) {
  VI_HIP(hipMemcpy2DAsync(dst, h->ldx * h->esz, src, h->g.nx * h->esz, h->g.nx * h->esz,
                          h->g.nz, hipMemcpyHostToDevice, h->stream));
  return 0;
}
int download_plane(vi_sim_s* h, void* dst, const void* src) {
  VI_HIP(hipMemcpy2DAsync(dst, h->g.nx * h->esz, src, h->ldx * h->esz, h->g.nx * h->esz,
                          h->g.nz, hipMemcpyDeviceToHost, h->stream));
  return 0;
}

template <typename R>
vi::StepParams<R> make_params(vi_sim_s* h) {
  vi::StepParams<R> p{};
  p.nx = h->g.nx;
  p.nz = h->g.nz;
  p.ldx = h->ldx;
  p.ntx = h->ntx;
  p.ntiles = h->ntiles;
  p.plane = h->plane;
  p.dt = (R)h->g.dt;
  double c[4];
  vi_fd_coeffs(h->g.dx, c);
  for (int l = 0; l < 4; ++l) p.Dx[l] = (R)c[l];
  vi_fd_coeffs(h->g.dz, c);
  for (int l = 0; l < 4; ++l) p.Dz[l] = (R)c[l];
  p.vx = (R*)h->state[0].p;
  p.vz = (R*)h->state[1].p;
  p.sxx = (R*)h->state[2].p;
  p.szz = (R*)h->state[3].p;
  p.sxz = (R*)h->state[4].p;
  p.M = (R*)h->M.p;
  p.rho = (const R*)h->model[0].p;
  p.lam = (const R*)h->model[1].p;
  p.mu = (const R*)h->model[2].p;
  p.tau = (const R*)h->model[3].p;
  for (int k = 0; k < h->L; ++k) {
    const R one = R(1), Lr = (R)h->L, tn = (R)h->taun[k];
    p.a1[k] = -one / (Lr * tn);  // visco-kernels.jl:50
    p.a2[k] = -one / tn;         // visco-kernels.jl:51
  }
  if (h->cpml_on) {
    p.T = h->cpml.thickness;
    p.xlo = h->cpml.left;
    p.xhi = h->cpml.right;
    p.zlo = h->cpml.top;
    p.zhi = h->cpml.bottom;
    R* cb = (R*)h->coef.p;
    const int nx = h->g.nx, nz = h->g.nz;
    p.ax = cb;
    p.bx = cb + nx;
    p.kx = cb + 2 * nx;
    p.az = cb + 3 * nx;
    p.bz = cb + 3 * nx + nz;
    p.kz = cb + 3 * nx + 2 * nz;
    for (int q = 0; q < 8; ++q) p.psi[q] = (R*)h->psi[q].p;
  }
  p.fs = h->fs;
  const int* st = (const int*)h->src_tab.p;
  p.src_off = st;
  p.src_cell = st + (h->ntiles + 1);
  p.src_kind = st + (h->ntiles + 1) + h->nsrc;
  p.src_row = st + (h->ntiles + 1) + 2 * h->nsrc;
  p.src_w = (const R*)h->src_w.p;
  p.snt = h->snt;
  const int* rt = (const int*)h->rec_tab.p;
  p.rec_off = rt;
  p.rec_cell = rt + (h->ntiles + 1);
  p.rec_comp = rt + (h->ntiles + 1) + h->nrec;
  p.rec_id = rt + (h->ntiles + 1) + 2 * h->nrec;
  p.traces = (R*)h->traces.p;
  p.nrec = h->nrec;
  p.rec_nt = h->rec_nt;
  p.step = h->step;
  return p;
}

int tile_of(const vi_sim_s* h, int ix, int iz) {
  return ((iz - 4) / (vi::NW * h->rz)) * h->ntx + (ix - 4) / vi::TX;
}

// Build a per-tile CSR table [off(ntiles+1) | cell(n) | a(n) | b(n)] and upload it.
int upload_table(vi_sim_s* h, DevBuf& buf, int n, const int* ix, const int* iz,
                 const int* a, const int* b) {
  std::vector<int> order(n);
  for (int q = 0; q < n; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return tile_of(h, ix[x], iz[x]) < tile_of(h, ix[y], iz[y]);
  });
  std::vector<int> tab((size_t)(h->ntiles + 1) + 3 * (size_t)n, 0);
  int* off = tab.data();
  int* cell = off + h->ntiles + 1;
  int* ca = cell + n;
  int* cb = ca + n;
  for (int q = 0; q < n; ++q) off[tile_of(h, ix[q], iz[q]) + 1]++;
  for (int t = 0; t < h->ntiles; ++t) off[t + 1] += off[t];
  for (int k = 0; k < n; ++k) {
    const int q = order[k];
    cell[k] = iz[q] * h->ldx + ix[q];
    ca[k] = a[q];
    cb[k] = b ? b[q] : q;
  }
  if (int rc = buf.alloc(tab.size() * sizeof(int), h->stream)) return rc;
  VI_HIP(hipMemcpyAsync(buf.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice,
                        h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

bool interior(const vi_sim_s* h, int ix, int iz) {
  return ix >= 4 && ix < h->g.nx - 4 && iz >= 4 && iz < h->g.nz - 4;
}

template <typename R, int RZ>
int launch_kernel(vi_sim_s* h, int k, const vi::StepParams<R>& p) {
  const dim3 block(vi::TX, vi::NW);
  const dim3 grid(h->ntiles);
  if (k == 0) {
    hipLaunchKernelGGL((vi::k_velocity<R, RZ>), grid, block, 0, h->stream, p);
  } else {
    switch (h->L) {
      case 0: hipLaunchKernelGGL((vi::k_stress<R, 0, RZ>), grid, block, 0, h->stream, p); break;
      case 1: hipLaunchKernelGGL((vi::k_stress<R, 1, RZ>), grid, block, 0, h->stream, p); break;
      case 2: hipLaunchKernelGGL((vi::k_stress<R, 2, RZ>), grid, block, 0, h->stream, p); break;
      case 3: hipLaunchKernelGGL((vi::k_stress<R, 3, RZ>), grid, block, 0, h->stream, p); break;
      case 4: hipLaunchKernelGGL((vi::k_stress<R, 4, RZ>), grid, block, 0, h->stream, p); break;
      case 5: hipLaunchKernelGGL((vi::k_stress<R, 5, RZ>), grid, block, 0, h->stream, p); break;
      case 6: hipLaunchKernelGGL((vi::k_stress<R, 6, RZ>), grid, block, 0, h->stream, p); break;
      default: return fail(VI_EINVAL, "nrelax out of range");
    }
  }
  VI_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int launch_step(vi_sim_s* h) {
  vi::StepParams<R> p = make_params<R>(h);
  for (int k = 0; k < 2; ++k) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (h->prof) {
      if (h->ev_used + 2 > h->ev.size()) {
        for (int q = 0; q < 2; ++q) {
          hipEvent_t e;
          VI_HIP(hipEventCreate(&e));
          h->ev.push_back(e);
          h->ev_kind.push_back(0);
        }
      }
      e0 = h->ev[h->ev_used];
      e1 = h->ev[h->ev_used + 1];
      h->ev_kind[h->ev_used] = k;
      h->ev_used += 2;
      VI_HIP(hipEventRecord(e0, h->stream));
    }
    int rc = 0;
    switch (h->rz) {
      case 1: rc = launch_kernel<R, 1>(h, k, p); break;
      case 2: rc = launch_kernel<R, 2>(h, k, p); break;
      case 4: rc = launch_kernel<R, 4>(h, k, p); break;
      case 8: rc = launch_kernel<R, 8>(h, k, p); break;
      default: return fail(VI_EINVAL, "bad tiling");
    }
    if (rc) return rc;
    if (h->prof) VI_HIP(hipEventRecord(e1, h->stream));
  }
  return 0;
}

// Re-tile the interior for rows-per-wave rz and rebuild the per-tile tables.
int apply_tiling(vi_sim_s* h, int rz) {
  h->rz = rz;
  h->ntx = (h->g.nx - 8 + vi::TX - 1) / vi::TX;
  h->ntz = (h->g.nz - 8 + vi::NW * rz - 1) / (vi::NW * rz);
  h->ntiles = h->ntx * h->ntz;
  if (int rc = upload_table(h, h->src_tab, h->nsrc, h->src_ix.data(), h->src_iz.data(),
                            h->src_kind.data(), nullptr))
    return rc;
  if (int rc = upload_table(h, h->rec_tab, h->nrec, h->rec_ix.data(), h->rec_iz.data(),
                            h->rec_comp.data(), nullptr))
    return rc;
  return 0;
}

// Default rows per wave: the largest variant that still gives about 16 waves per CU.
int default_rz(const vi_grid_desc& g) {
  const long cols = (g.nx - 8 + vi::TX - 1) / vi::TX;
  for (int rz : {8, 4, 2}) {
    const long waves = cols * ((g.nz - 8 + rz - 1) / rz);
    if (waves >= 256L * 16) return rz;
  }
  return 1;
}

int collect_profile(vi_sim_s* h) {
  if (!h->ev_used) return 0;
  VI_HIP(hipStreamSynchronize(h->stream));
  for (size_t q = 0; q + 1 < h->ev_used; q += 2) {
    float ms = 0.f;
    VI_HIP(hipEventElapsedTime(&ms, h->ev[q], h->ev[q + 1]));
    h->prof_ms[h->ev_kind[q]] += ms;
    h->prof_n[h->ev_kind[q]] += 1;
  }
  h->ev_used = 0;
  return 0;
}

size_t field_count(const vi_sim_s* h, int field) {
  const size_t nc = (size_t)h->g.nx * h->g.nz;
  if (field >= VI_VX && field <= VI_SXZ) return nc;
  if (field >= VI_MXX && field <= VI_MXZ) return h->L ? (size_t)h->L * nc : 0;
  if (field >= VI_PSI_VX_X && field <= VI_PSI_SZZ_Z) {
    if (!h->cpml_on) return 0;
    const int T = h->cpml.thickness;
    return field < VI_PSI_VX_Z ? (size_t)h->g.nz * 2 * T : (size_t)2 * T * h->g.nx;
  }
  return 0;
}

}  // namespace

extern "C" {

const char* vi_last_error(void) { return g_err.c_str(); }
const char* vi_version(void) { return "vi_mi355x 0.1 (gfx950)"; }

void vi_fd_coeffs(double h, double out[4]) {
  // _stencil(x, 0, 1) (Stencil.jl:9-14) at x = +-1/2 .. +-7/2, folded to the half stencil
  // of diff-funcs.jl:3-32.
  out[0] = (1225.0 / 1024.0) / h;
  out[1] = (-245.0 / 3072.0) / h;
  out[2] = (49.0 / 5120.0) / h;
  out[3] = (-5.0 / 7168.0) / h;
}

double vi_cpml_auto_d0(int thickness, double vmax, double dx, double R, double power) {
  return -(power + 1.0) * vmax * std::log(R) / (2.0 * thickness * dx);  // CPML.jl:60
}

int vi_cpml_profile(int n, const vi_cpml_desc* c, double dt, int axis, double* a, double* b,
                    double* kappa) {
  if (!c || n <= 0 || !a || !b || !kappa || c->thickness < 0 || 2 * c->thickness > n)
    return fail(VI_EINVAL, "vi_cpml_profile: bad arguments");
  const int lo = axis == 0 ? c->left : c->top;
  const int hi = axis == 0 ? c->right : c->bottom;
  cpml_axis(n, c->thickness, c->d0, c->alpha_max, c->kappa_max, c->power, dt, lo, hi, a, b,
            kappa);
  return 0;
}

int vi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int vi_create(const vi_grid_desc* g, vi_handle* out) {
  if (!g || !out) return fail(VI_EINVAL, "vi_create: null argument");
  *out = nullptr;
  if (g->nx < 16 || g->nz < 16) return fail(VI_EINVAL, "vi_create: nx and nz must be >= 16");
  if (g->precision != VI_F32 && g->precision != VI_F64)
    return fail(VI_EINVAL, "vi_create: precision must be VI_F32 or VI_F64");
  if (g->nrelax < 0 || g->nrelax > vi::MAXL)
    return fail(VI_EINVAL, "vi_create: nrelax must be in [0, 6]");
  if (g->fd_order_x != 8 || g->fd_order_z != 8)
    return fail(VI_EINVAL, "vi_create: only 8th-order staggered operators are implemented");
  if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0))
    return fail(VI_EINVAL, "vi_create: dx, dz, dt must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(VI_EDEVICE, "vi_create: no HIP device visible");
  std::unique_ptr<vi_sim_s> h(new vi_sim_s());
  h->g = *g;
  h->L = g->nrelax;
  h->esz = g->precision == VI_F32 ? 4 : 8;
  VI_HIP(hipGetDevice(&h->device));
  VI_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->ldx = (int)round_up((size_t)g->nx + 72, 64);
  h->nzp = g->nz + vi::NW * vi::RZ_MAX + 16;
  h->plane = (size_t)h->ldx * h->nzp;
  const size_t pb = h->plane * h->esz;
  for (auto& b : h->state)
    if (int rc = b.alloc(pb, h->stream)) return rc;
  if (h->L)
    if (int rc = h->M.alloc(3 * (size_t)h->L * pb, h->stream)) return rc;
  for (int q = 0; q < 4; ++q)
    if (int rc = h->model[q].alloc(pb, h->stream)) return rc;
  // tiling and the (empty) source / receiver tables
  if (int rc = apply_tiling(h.get(), default_rz(*g))) return rc;
  VI_HIP(hipStreamSynchronize(h->stream));
  *out = h.release();
  return 0;
}

int vi_destroy(vi_handle h) {
  if (!h) return 0;
  (void)hipStreamSynchronize(h->stream);
  for (auto e : h->ev) (void)hipEventDestroy(e);
  hipStream_t s = h->stream;
  delete h;
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int vi_set_model(vi_handle h, const void* rho, const void* lam, const void* mu,
                 const void* tau) {
  if (int rc = check_handle(h)) return rc;
  if (!rho || !lam || !mu) return fail(VI_EINVAL, "vi_set_model: rho, lam, mu required");
  if (h->L && !tau) return fail(VI_EINVAL, "vi_set_model: tau required when nrelax > 0");
  const void* src[4] = {rho, lam, mu, tau};
  for (int q = 0; q < 4; ++q)
    if (src[q])
      if (int rc = upload_plane(h, h->model[q].p, src[q])) return rc;
  VI_HIP(hipStreamSynchronize(h->stream));
  h->model_set = true;
  return 0;
}

int vi_set_relaxation(vi_handle h, const double* taun) {
  if (int rc = check_handle(h)) return rc;
  if (h->L == 0) return 0;
  if (!taun) return fail(VI_EINVAL, "vi_set_relaxation: null taun");
  for (int k = 0; k < h->L; ++k) {
    if (!(taun[k] > 0)) return fail(VI_EINVAL, "vi_set_relaxation: taun must be positive");
    h->taun[k] = taun[k];
  }
  h->relax_set = true;
  return 0;
}

int vi_set_cpml(vi_handle h, const vi_cpml_desc* c) {
  if (int rc = check_handle(h)) return rc;
  if (!c) {
    h->cpml_on = false;
    return 0;
  }
  if (c->thickness < 1 || 2 * c->thickness > h->g.nx || 2 * c->thickness > h->g.nz)
    return fail(VI_EINVAL, "vi_set_cpml: thickness must be in [1, min(nx, nz)/2]");
  if (!(c->d0 >= 0) || !(c->kappa_max >= 1) || !(c->alpha_max >= 0) || !(c->power >= 0))
    return fail(VI_EINVAL, "vi_set_cpml: need d0 >= 0, kappa_max >= 1, alpha_max >= 0");
  if (c->top && h->fs)
    return fail(VI_EINVAL, "vi_set_cpml: top strip conflicts with the free surface");
  const int nx = h->g.nx, nz = h->g.nz, T = c->thickness;
  std::vector<double> a(std::max(nx, nz)), b(std::max(nx, nz)), k(std::max(nx, nz));
  std::vector<char> hostc((size_t)(3 * nx + 3 * nz) * h->esz);
  auto put = [&](size_t off, const std::vector<double>& v, int n) {
    for (int q = 0; q < n; ++q) {
      if (h->esz == 4) ((float*)hostc.data())[off + q] = (float)v[q];
      else ((double*)hostc.data())[off + q] = v[q];
    }
  };
  cpml_axis(nx, T, c->d0, c->alpha_max, c->kappa_max, c->power, h->g.dt, c->left, c->right,
            a.data(), b.data(), k.data());
  put(0, a, nx);
  put(nx, b, nx);
  put(2 * nx, k, nx);
  cpml_axis(nz, T, c->d0, c->alpha_max, c->kappa_max, c->power, h->g.dt, c->top, c->bottom,
            a.data(), b.data(), k.data());
  put(3 * nx, a, nz);
  put(3 * nx + nz, b, nz);
  put(3 * nx + 2 * nz, k, nz);
  if (int rc = h->coef.alloc(hostc.size(), h->stream)) return rc;
  VI_HIP(hipMemcpyAsync(h->coef.p, hostc.data(), hostc.size(), hipMemcpyHostToDevice,
                        h->stream));
  for (int q = 0; q < 4; ++q)
    if (int rc = h->psi[q].alloc((size_t)nz * 2 * T * h->esz, h->stream)) return rc;
  for (int q = 4; q < 8; ++q)
    if (int rc = h->psi[q].alloc((size_t)2 * T * h->ldx * h->esz, h->stream)) return rc;
  VI_HIP(hipStreamSynchronize(h->stream));
  h->cpml = *c;
  h->cpml_on = true;
  return 0;
}

int vi_set_free_surface(vi_handle h, int enabled) {
  if (int rc = check_handle(h)) return rc;
  if (enabled && h->cpml_on && h->cpml.top)
    return fail(VI_EINVAL, "vi_set_free_surface: conflicts with a top CPML strip");
  if (enabled && h->g.nz < 17) return fail(VI_EINVAL, "vi_set_free_surface: nz must be >= 17");
  h->fs = enabled ? 1 : 0;
  return 0;
}

int vi_set_sources(vi_handle h, int n, const int* ix, const int* iz, const int* kind, int nt,
                   const void* wavelets) {
  if (int rc = check_handle(h)) return rc;
  if (n < 0 || nt < 0 || (n > 0 && (!ix || !iz || !kind || (nt > 0 && !wavelets))))
    return fail(VI_EINVAL, "vi_set_sources: bad arguments");
  for (int q = 0; q < n; ++q) {
    if (!interior(h, ix[q], iz[q]))
      return fail(VI_EINVAL, "vi_set_sources: source " + std::to_string(q) +
                                 " is outside the interior [4, nx-4) x [4, nz-4)");
    if (kind[q] < 0 || kind[q] > 2) return fail(VI_EINVAL, "vi_set_sources: bad kind");
  }
  h->src_ix.assign(ix, ix + n);
  h->src_iz.assign(iz, iz + n);
  h->src_kind.assign(kind, kind + n);
  h->nsrc = n;
  if (int rc = upload_table(h, h->src_tab, n, ix, iz, kind, nullptr)) return rc;
  if (int rc = h->src_w.alloc((size_t)n * nt * h->esz, h->stream)) return rc;
  if (n && nt)
    VI_HIP(hipMemcpyAsync(h->src_w.p, wavelets, (size_t)n * nt * h->esz,
                          hipMemcpyHostToDevice, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->nsrc = n;
  h->snt = nt;
  return 0;
}

int vi_set_receivers(vi_handle h, int n, const int* ix, const int* iz, const int* comp,
                     int nt_max) {
  if (int rc = check_handle(h)) return rc;
  if (n < 0 || nt_max < 0 || (n > 0 && (!ix || !iz || !comp)))
    return fail(VI_EINVAL, "vi_set_receivers: bad arguments");
  for (int q = 0; q < n; ++q) {
    if (!interior(h, ix[q], iz[q]))
      return fail(VI_EINVAL, "vi_set_receivers: receiver " + std::to_string(q) +
                                 " is outside the interior [4, nx-4) x [4, nz-4)");
    if (comp[q] < 0 || comp[q] > 4) return fail(VI_EINVAL, "vi_set_receivers: bad component");
  }
  h->rec_ix.assign(ix, ix + n);
  h->rec_iz.assign(iz, iz + n);
  h->rec_comp.assign(comp, comp + n);
  h->nrec = n;
  if (int rc = upload_table(h, h->rec_tab, n, ix, iz, comp, nullptr)) return rc;
  if (int rc = h->traces.alloc((size_t)n * nt_max * h->esz, h->stream)) return rc;
  VI_HIP(hipStreamSynchronize(h->stream));
  h->nrec = n;
  h->rec_nt = nt_max;
  return 0;
}

int vi_set_state(vi_handle h, int field, const void* host) {
  if (int rc = check_handle(h)) return rc;
  if (!host) return fail(VI_EINVAL, "vi_set_state: null buffer");
  const size_t n = field_count(h, field);
  if (!n) return fail(VI_EINVAL, "vi_set_state: field not present in this configuration");
  if (field <= VI_SXZ) {
    if (int rc = upload_plane(h, h->state[field].p, host)) return rc;
  } else if (field <= VI_MXZ) {
    const size_t nc = (size_t)h->g.nx * h->g.nz;
    for (int k = 0; k < h->L; ++k) {
      char* dst = (char*)h->M.p + ((size_t)(field - VI_MXX) * h->L + k) * h->plane * h->esz;
      if (int rc = upload_plane(h, dst, (const char*)host + k * nc * h->esz)) return rc;
    }
  } else {
    const int q = field - VI_PSI_VX_X;
    if (q < 4) {
      VI_HIP(hipMemcpyAsync(h->psi[q].p, host, n * h->esz, hipMemcpyHostToDevice, h->stream));
    } else {
      const int T2 = 2 * h->cpml.thickness;
      VI_HIP(hipMemcpy2DAsync(h->psi[q].p, h->ldx * h->esz, host, h->g.nx * h->esz,
                              h->g.nx * h->esz, T2, hipMemcpyHostToDevice, h->stream));
    }
  }
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int vi_get_state(vi_handle h, int field, void* host) {
  if (int rc = check_handle(h)) return rc;
  if (!host) return fail(VI_EINVAL, "vi_get_state: null buffer");
  const size_t n = field_count(h, field);
  if (!n) return fail(VI_EINVAL, "vi_get_state: field not present in this configuration");
  if (field <= VI_SXZ) {
    if (int rc = download_plane(h, host, h->state[field].p)) return rc;
  } else if (field <= VI_MXZ) {
    const size_t nc = (size_t)h->g.nx * h->g.nz;
    for (int k = 0; k < h->L; ++k) {
      const char* src =
          (const char*)h->M.p + ((size_t)(field - VI_MXX) * h->L + k) * h->plane * h->esz;
      if (int rc = download_plane(h, (char*)host + k * nc * h->esz, src)) return rc;
    }
  } else {
    const int q = field - VI_PSI_VX_X;
    if (q < 4) {
      VI_HIP(hipMemcpyAsync(host, h->psi[q].p, n * h->esz, hipMemcpyDeviceToHost, h->stream));
    } else {
      const int T2 = 2 * h->cpml.thickness;
      VI_HIP(hipMemcpy2DAsync(host, h->g.nx * h->esz, h->psi[q].p, h->ldx * h->esz,
                              h->g.nx * h->esz, T2, hipMemcpyDeviceToHost, h->stream));
    }
  }
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int vi_reset(vi_handle h) {
  if (int rc = check_handle(h)) return rc;
  for (auto& b : h->state)
    if (b.p) VI_HIP(hipMemsetAsync(b.p, 0, b.bytes, h->stream));
  if (h->M.p) VI_HIP(hipMemsetAsync(h->M.p, 0, h->M.bytes, h->stream));
  for (auto& b : h->psi)
    if (b.p) VI_HIP(hipMemsetAsync(b.p, 0, b.bytes, h->stream));
  if (h->traces.p) VI_HIP(hipMemsetAsync(h->traces.p, 0, h->traces.bytes, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->step = 0;
  return 0;
}

int vi_step(vi_handle h, int nsteps) {
  if (int rc = check_handle(h)) return rc;
  if (nsteps < 0) return fail(VI_EINVAL, "vi_step: nsteps < 0");
  if (!h->model_set) return fail(VI_ESTATE, "vi_step: vi_set_model has not been called");
  if (h->L && !h->relax_set)
    return fail(VI_ESTATE, "vi_step: vi_set_relaxation has not been called");
  for (int k = 0; k < nsteps; ++k) {
    int rc = h->esz == 4 ? launch_step<float>(h) : launch_step<double>(h);
    if (rc) return rc;
    h->step++;
  }
  return 0;
}

int vi_sync(vi_handle h) {
  if (int rc = check_handle(h)) return rc;
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

long vi_step_index(vi_handle h) { return h ? h->step : -1; }

int vi_get_traces(vi_handle h, void* out, int nt) {
  if (int rc = check_handle(h)) return rc;
  if (nt < 0 || nt > h->rec_nt || (!out && nt && h->nrec))
    return fail(VI_EINVAL, "vi_get_traces: nt exceeds the recorded capacity");
  if (!h->nrec || !nt) return 0;
  std::vector<char> tmp((size_t)nt * h->nrec * h->esz);
  VI_HIP(hipMemcpyAsync(tmp.data(), h->traces.p, tmp.size(), hipMemcpyDeviceToHost, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  // device layout [t][rec] -> host [rec][t]
  for (int r = 0; r < h->nrec; ++r)
    for (int t = 0; t < nt; ++t)
      std::memcpy((char*)out + ((size_t)r * nt + t) * h->esz,
                  tmp.data() + ((size_t)t * h->nrec + r) * h->esz, h->esz);
  return 0;
}

int vi_profile_enable(vi_handle h, int enabled) {
  if (int rc = check_handle(h)) return rc;
  if (int rc = collect_profile(h)) return rc;
  h->prof = enabled != 0;
  h->prof_ms[0] = h->prof_ms[1] = 0;
  h->prof_n[0] = h->prof_n[1] = 0;
  return 0;
}

int vi_profile_read(vi_handle h, int kernel, double* ms_total, long* launches) {
  if (int rc = check_handle(h)) return rc;
  if (kernel < 0 || kernel > 1 || !ms_total || !launches)
    return fail(VI_EINVAL, "vi_profile_read: bad arguments");
  if (int rc = collect_profile(h)) return rc;
  *ms_total = h->prof_ms[kernel];
  *launches = h->prof_n[kernel];
  return 0;
}

int vi_algorithmic_bytes(vi_handle h, int kernel, double* bytes) {
  if (int rc = check_handle(h)) return rc;
  if (kernel < 0 || kernel > 1 || !bytes) return fail(VI_EINVAL, "vi_algorithmic_bytes");
  const double ni = (double)(h->g.nx - 8) * (h->g.nz - 8);
  // velocity: v R+W (4) + s R (3) + rho (1); stress: v R (2) + s R+W (6) + lam, mu (2)
  // [+ tau (1) + M R+W (6L)]
  double words = kernel == 0 ? 8.0 : (h->L ? 11.0 + 6.0 * h->L : 10.0);
  double total = words * ni;
  if (h->cpml_on) {  // two psi arrays read and written per strip direction per kernel
    const int T = h->cpml.thickness;
    long xcols = 0, zrows = 0;
    for (int i = 4; i < h->g.nx - 4; ++i)
      xcols += (h->cpml.left && i < T) || (h->cpml.right && i >= h->g.nx - T);
    for (int j = 4; j < h->g.nz - 4; ++j)
      zrows += (h->cpml.top && j < T) || (h->cpml.bottom && j >= h->g.nz - T);
    total += 4.0 * xcols * (h->g.nz - 8) + 4.0 * zrows * (h->g.nx - 8);
  }
  *bytes = total * (double)h->esz;
  return 0;
}

int vi_set_tiling(vi_handle h, int rz) {
  if (int rc = check_handle(h)) return rc;
  if (rz != 1 && rz != 2 && rz != 4 && rz != 8)
    return fail(VI_EINVAL, "vi_set_tiling: rows per wavefront must be 1, 2, 4 or 8");
  VI_HIP(hipStreamSynchronize(h->stream));
  return apply_tiling(h, rz);
}

int vi_get_tiling(vi_handle h) { return h ? h->rz : -1; }

int vi_autotune(vi_handle h, int trial_steps, int* chosen, double* us_per_step) {
  if (int rc = check_handle(h)) return rc;
  if (trial_steps < 1) return fail(VI_EINVAL, "vi_autotune: trial_steps must be >= 1");
  if (!h->model_set || (h->L && !h->relax_set))
    return fail(VI_ESTATE, "vi_autotune: configure the model first");
  VI_HIP(hipStreamSynchronize(h->stream));
  // snapshot every mutable array so that the trials leave the state untouched
  std::vector<DevBuf*> live;
  for (auto& b : h->state) live.push_back(&b);
  live.push_back(&h->M);
  for (auto& b : h->psi) live.push_back(&b);
  live.push_back(&h->traces);
  std::vector<DevBuf> snap(live.size());
  for (size_t q = 0; q < live.size(); ++q) {
    if (!live[q]->p) continue;
    if (int rc = snap[q].alloc(live[q]->bytes, h->stream)) return rc;
    VI_HIP(hipMemcpyAsync(snap[q].p, live[q]->p, live[q]->bytes, hipMemcpyDeviceToDevice,
                          h->stream));
  }
  auto restore = [&]() -> int {
    for (size_t q = 0; q < live.size(); ++q)
      if (live[q]->p)
        VI_HIP(hipMemcpyAsync(live[q]->p, snap[q].p, live[q]->bytes, hipMemcpyDeviceToDevice,
                              h->stream));
    return 0;
  };
  const long step0 = h->step;
  const bool prof0 = h->prof;
  h->prof = false;
  hipEvent_t e0, e1;
  VI_HIP(hipEventCreate(&e0));
  VI_HIP(hipEventCreate(&e1));
  const int cands[4] = {1, 2, 4, 8};
  int best = h->rz;
  double best_us = 1e300;
  int rc = 0;
  for (int c = 0; c < 4 && !rc; ++c) {
    if ((rc = apply_tiling(h, cands[c]))) break;
    for (int k = 0; k < 2 && !rc; ++k) rc = h->esz == 4 ? launch_step<float>(h) : launch_step<double>(h);
    if (rc) break;
    if ((rc = hipEventRecord(e0, h->stream) == hipSuccess ? 0 : VI_EDEVICE)) break;
    for (int k = 0; k < trial_steps && !rc; ++k) {
      rc = h->esz == 4 ? launch_step<float>(h) : launch_step<double>(h);
      h->step++;
    }
    if (rc) break;
    if ((rc = hipEventRecord(e1, h->stream) == hipSuccess ? 0 : VI_EDEVICE)) break;
    if ((rc = hipEventSynchronize(e1) == hipSuccess ? 0 : VI_EDEVICE)) break;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double us = 1e3 * ms / trial_steps;
    if (us_per_step) us_per_step[c] = us;
    if (us < best_us) {
      best_us = us;
      best = cands[c];
    }
    h->step = step0;
    if ((rc = restore())) break;
  }
  h->step = step0;
  h->prof = prof0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    (void)restore();
    (void)hipStreamSynchronize(h->stream);
    return rc == VI_EDEVICE && g_err.empty() ? fail(rc, "vi_autotune: HIP error") : rc;
  }
  VI_HIP(hipStreamSynchronize(h->stream));
  if (int r2 = apply_tiling(h, best)) return r2;
  if (chosen) *chosen = best;
  return 0;
}

}  // extern "C"

