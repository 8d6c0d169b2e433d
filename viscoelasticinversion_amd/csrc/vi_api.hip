// vi_api.hip -- host side of the C ABI declared in include/vi_mi355x.h.
//
// vi_sim_s (a handle) owns the device state of one shot in the padded layout the kernels
// expect, a reference to a device-resident model (its own or a column window of a shared
// vi_model), one HIP stream, the per-tile source/receiver tables and a device ShotDesc.
// vi_batch_s advances a set of handles with one launch per half step (per precision and L
// group), optionally replaying hipGraphs of a block of steps.
#include "vi_mi355x.h"
#include "vi_kernels.hpp"
#include "vi_adjoint.hpp"
#include "vi_rates.hpp"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
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

#define VI_TRY(expr)     \
  do {                   \
    int rc_ = (expr);    \
    if (rc_) return rc_; \
  } while (0)

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// CPMLCoefficients along one axis (CPML.jl:109-189) in the working precision F, as the
// reference computes them for its F arrays: the profiles d, alpha, kappa are evaluated in
// double and stored in F (CPML.jl:91-161); where d > 0 or alpha > 0, b = F(exp(-(d/kappa +
// alpha) dt)) with d/kappa + alpha summed in F and the product with the double dt and the
// exponential in double (CPML.jl:172-174), and a = d/(kappa (d + kappa alpha)) (b - 1) in F
// where |d| > eps(F) (CPML.jl:175-177).
template <typename F>
void cpml_axis(int n, int T, double d0, double alpha_max, double kappa_max, double power,
               double dt, int lo, int hi, F* a, F* b, F* kappa) {
  for (int i = 0; i < n; ++i) {
    double p = -1.0;
    if (lo && i < T) p = (T - i - 0.5) / T;             // CPML.jl:124
    if (hi && i >= n - T) p = (i - (n - T) + 0.5) / T;  // CPML.jl:134
    F d = 0, alpha = 0, k = 1;
    if (p >= 0.0) {
      d = (F)(d0 * std::pow(p, power));          // CPML.jl:91-93
      alpha = (F)(alpha_max * (1.0 - p));        // CPML.jl:95-98
      k = (F)(1.0 + (kappa_max - 1.0) * p);      // CPML.jl:100-102
    }
    F bb = 0, aa = 0;
    if (d > 0 || alpha > 0) {
      const F dk = d / k;
      const F sum = dk + alpha;
      bb = (F)std::exp(-(double)sum * dt);                                            // :174
      if (std::fabs(d) > std::numeric_limits<F>::epsilon())
        aa = d / (k * (d + k * alpha)) * (bb - F(1));                                   // :176
    }
    a[i] = aa;
    b[i] = bb;
    kappa[i] = k;
  }
}

// Pinned host memory, grown on demand.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  int reserve(size_t nbytes);
};

// Large device buffers released by any handle are kept for the next large allocation on the
// same device that they fit (at most a quarter larger than asked), instead of going back to
// the runtime. Jobs run one after another with fresh handles (the serving loop, gradients
// shot by shot) otherwise hand tens of GB back and forth, and such a hipMalloc or hipFree
// sporadically took 0.1-1.5 s on MI355X (50-60 GB gradient workspaces). A kept buffer is
// idle: its device is synchronised when the buffer is released, as hipFree would. At most a
// third of the device's memory is kept (the oldest buffers are freed first); when an
// allocation fails, every kept buffer of that device is freed and the allocation is tried
// once more. vi_mem_info counts kept buffers as free; vi_release_cached frees them.
namespace spare {
constexpr size_t MIN_BYTES = size_t(64) << 20;
struct Block {
  int device;
  size_t cap;
  void* p;
};
struct Cache {
  std::mutex mu;
  std::vector<Block> blocks;  // oldest first
};
// never destroyed: nothing is handed back to the runtime during process exit
Cache& cache() {
  static Cache* c = new Cache;
  return *c;
}

size_t bytes_on(int device) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  size_t t = 0;
  for (const Block& b : c.blocks)
    if (b.device == device) t += b.cap;
  return t;
}

// The smallest kept buffer of `device` that holds n bytes and is at most n + n/4 (false: none).
bool take(int device, size_t n, void** p, size_t* cap) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  size_t best = c.blocks.size();
  for (size_t q = 0; q < c.blocks.size(); ++q) {
    const Block& b = c.blocks[q];
    if (b.device == device && b.cap >= n && b.cap - n <= n / 4 &&
        (best == c.blocks.size() || b.cap < c.blocks[best].cap))
      best = q;
  }
  if (best == c.blocks.size()) return false;
  *p = c.blocks[best].p;
  *cap = c.blocks[best].cap;
  c.blocks.erase(c.blocks.begin() + (long)best);
  return true;
}

// Free every kept buffer of `device`; true when there was one.
bool drop(int device) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  bool any = false;
  for (size_t q = 0; q < c.blocks.size();) {
    if (c.blocks[q].device == device) {
      (void)hipFree(c.blocks[q].p);
      c.blocks.erase(c.blocks.begin() + (long)q);
      any = true;
    } else {
      ++q;
    }
  }
  return any;
}

// Release a buffer: small ones to the runtime, large ones to the cache once the device is idle.
void give(int device, size_t cap, void* p) {
  if (cap < MIN_BYTES) {
    (void)hipFree(p);
    return;
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != device) (void)hipSetDevice(device);
  const bool idle = hipDeviceSynchronize() == hipSuccess;
  size_t fr = 0, total = 0;
  (void)hipMemGetInfo(&fr, &total);
  if (cur != device) (void)hipSetDevice(cur);
  if (!idle) {  // a failed device: nothing is kept
    (void)hipFree(p);
    return;
  }
  Cache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  c.blocks.push_back({device, cap, p});
  size_t kept = 0;
  for (const Block& b : c.blocks)
    if (b.device == device) kept += b.cap;
  for (size_t q = 0; q < c.blocks.size() && kept > total / 3;) {
    if (c.blocks[q].device == device) {
      kept -= c.blocks[q].cap;
      (void)hipFree(c.blocks[q].p);
      c.blocks.erase(c.blocks.begin() + (long)q);
    } else {
      ++q;
    }
  }
}
}  // namespace spare

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;  // as asked
  size_t cap = 0;    // as allocated (a reused buffer may be larger)
  int device = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) spare::give(device, cap, p);
    p = nullptr;
    bytes = cap = 0;
  }
  // p = a kept buffer of the current device, else a new allocation (retried once after
  // freeing the kept buffers when the device is out of memory)
  int obtain(size_t nbytes) {
    VI_HIP(hipGetDevice(&device));
    if (nbytes >= spare::MIN_BYTES && spare::take(device, nbytes, &p, &cap)) return 0;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      if (spare::drop(device)) e = hipMalloc(&p, nbytes);
    }
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? VI_ENOMEM : VI_EDEVICE,
                  "hipMalloc(" + std::to_string(nbytes) + " bytes): " + hipGetErrorString(e));
    }
    cap = nbytes;
    return 0;
  }
  int alloc(size_t nbytes, hipStream_t s) {
    release();
    if (!nbytes) return 0;
    VI_TRY(obtain(nbytes));
    bytes = nbytes;
    VI_HIP(hipMemsetAsync(p, 0, nbytes, s));
    return 0;
  }
  // grow-only allocation without clearing (buffers that are written before they are read)
  int reserve(size_t nbytes) {
    if (nbytes <= bytes && p) return 0;
    release();
    if (!nbytes) return 0;
    VI_TRY(obtain(nbytes));
    bytes = nbytes;
    return 0;
  }
};

int HostBuf::reserve(size_t nbytes) {
  if (nbytes <= bytes && p) return 0;
  release();
  if (!nbytes) return 0;
  VI_HIP(hipHostMalloc(&p, nbytes, 0));
  bytes = nbytes;
  return 0;
}

// HIP events around every launch of one stream, summed per kernel kind.
struct Prof {
  bool on = false;
  int every = 1;   // time the launches of every `every`-th step
  long count = 0;  // steps seen since enable, counted across calls
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
  size_t used = 0;
  double ms[2] = {0, 0};
  long n[2] = {0, 0};
  ~Prof() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  // An event pair that the launch itself records (hipExtLaunchKernelGGL), so the timing is
  // the kernel's execution without extra marker packets in the stream.
  int begin(int k, hipEvent_t* e0, hipEvent_t* e1) {
    if (used + 2 > ev.size()) {
      for (int q = 0; q < 2; ++q) {
        hipEvent_t e;
        VI_HIP(hipEventCreate(&e));
        ev.push_back(e);
        kind.push_back(0);
      }
    }
    kind[used] = k;
    *e0 = ev[used];
    *e1 = ev[used + 1];
    used += 2;
    return 0;
  }
  int collect(hipStream_t s) {
    if (!used) return 0;
    VI_HIP(hipStreamSynchronize(s));
    for (size_t q = 0; q + 1 < used; q += 2) {
      float t = 0.f;
      VI_HIP(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
      ms[kind[q]] += t;
      n[kind[q]] += 1;
    }
    used = 0;
    return 0;
  }
  void reset() {
    ms[0] = ms[1] = 0;
    n[0] = n[1] = 0;
    count = 0;
  }
  // called once per step: whether this step's launches carry events
  bool tick() { return on && (count++ % every) == 0; }
};

}  // namespace

// A device-resident model: rho, lam, mu, tau planes with padded pitch.
struct ModelStore {
  int nx = 0, nz = 0, ldm = 0, nzp = 0;
  size_t esz = 4;
  bool has_tau = false;
  bool vti = false;
  DevBuf arr[6];  // rho, lam | C13, mu | C55, tau, C11, C33 (the last two for VTI only)
};

struct vi_model_s {
  std::shared_ptr<ModelStore> m;
};

struct vi_sim_s {
  vi_grid_desc g{};
  int L = 0;
  int ldx = 0, nzp = 0;
  size_t plane = 0;
  int ntx = 0, ntz = 0, ntiles = 0;
  int rz = 8;
  int device = 0;
  hipStream_t stream = nullptr;
  size_t esz = 4;
  DevBuf state[5];  // vx vz sxx szz sxz
  DevBuf M;
  std::shared_ptr<ModelStore> model;
  int model_x0 = 0;
  DevBuf psi[8];
  DevBuf coef;  // ax bx kx (nx each), az bz kz (nz each)
  DevBuf src_tab, src_w, rec_tab, traces;
  DevBuf src_wc;                   // the wavelets [snt][nsrc] in the CSR order of src_tab
  std::vector<char> src_w_host;    // the wavelets as given [nsrc][snt]
  double taun[vi::MAXL] = {};
  bool relax_set = false;
  bool cpml_on = false;
  vi_cpml_desc cpml{};
  int fs = 0;                 // 0 none, 1 stress imaging, 2 H-AFDA, 3 H-AFDA + formula #3
  int afda_pts[13] = {3, 4, 6, 8, 3, 4, 6, 2, 4, 6, 3, 4, 6};  // points per one-sided stencil
  int herm_pts = 5;           // half points of formula #3 (fs = 3)
  DevBuf afda;                // H-AFDA tables (see upload_afda) in the handle's precision
  int nsrc = 0, snt = 0;
  int nrec = 0, rec_nt = 0;
  std::vector<int> src_ix, src_iz, src_kind, rec_ix, rec_iz, rec_comp;
  long step = 0;
  // device descriptor for launches of this handle alone
  std::vector<char> desc_host;
  DevBuf desc_dev, tile_shot_dev;
  DevBuf rates_ws;  // vi_rates workspace (rate planes and the source list), grown on demand
  // gradient workspace kept between gradient calls (checkpoints, segment slots, adjoint
  // state), so that repeated gradients do not allocate and free it each time; vi_trim frees it
  DevBuf grad_ck, grad_seg, grad_adj;
  HostBuf grad_host;  // pinned staging of the gradient outputs
  DevBuf check_flag;  // vi_check_finite's flag
  bool desc_dirty = true;
  unsigned long version = 1;  // bumped on every change that invalidates a descriptor
  // ordering with batches
  hipEvent_t own_ev = nullptr, ext_ev = nullptr;
  bool own_pending = false, ext_pending = false;
  Prof prof;
  bool grad_busy = false;  // a gradient job owns the handle
};

struct BatchGroup {
  int esz = 4, L = 0;
  bool afda = false;  // some member has the H-AFDA surface (launch that variant)
  std::vector<vi_sim_s*> members;
  int ntiles = 0;
  std::vector<char> desc_host;
  DevBuf desc_dev, tile_shot_dev;
};

struct vi_batch_s {
  std::vector<vi_sim_s*> members;
  std::vector<unsigned long> versions;
  std::vector<long> steps;  // each member's step index as the batch descriptors hold it
  std::vector<std::unique_ptr<BatchGroup>> groups;
  bool built = false;
  // rows per wave of the members' tiles; 4-row tiles by default (the 64-shot cache-blocked
  // batch runs within 0.5 % of its fastest at 4- and 8-row tiles; vi_batch_autotune picks)
  int rz = 1;
  int graph_steps = 0;
  // cache blocking over shots (0: every shot in every launch); default: groups of at most
  // 224 MB of working set, which stay in the 256 MB Infinity Cache through a block of steps
  // (64 shots x U{512..1536} x 512, L = 3, FP32: 38.5 GPt/s against 32.4 in lockstep, and
  // 37.0 / 38.2 / 37.0 at 160 / 192 / 256 MB)
  size_t block_bytes = (size_t)224 << 20;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  hipStream_t stream = nullptr;
  // side streams: with n streams, consecutive launch groups take the streams in turn and n
  // groups (each at most 1/n of the cache block) run at once, overlapping each other's launch
  // ramp and drain; the side streams fork from stream before a block and join it after
  // (events)
  static constexpr int MAX_STREAMS = 4;
  int nstreams = 2;
  hipStream_t side[MAX_STREAMS - 1] = {};
  hipEvent_t fork_ev = nullptr, join_ev[MAX_STREAMS - 1] = {};
  Prof prof;
};

// A gradient computation in flight on one handle (see GradRun).
struct vi_grad_s {
  vi_sim_s* h = nullptr;
  hipEvent_t ev = nullptr;  // recorded after the last enqueued chunk
  bool done = false;
  virtual ~vi_grad_s() {
    if (ev) (void)hipEventDestroy(ev);
  }
  virtual int advance(int max_steps) = 0;
  virtual int finish(double* misfit, void* const out[6]) = 0;
};

namespace {

int check_handle(vi_handle h) {
  if (!h) return fail(VI_EINVAL, "null handle");
  if (h->grad_busy) return fail(VI_ESTATE, "handle in use by a gradient job");
  return 0;
}

void touch(vi_sim_s* h) {
  h->desc_dirty = true;
  h->version++;
}

// Make this handle's stream wait for batch work that touched its arrays.
int wait_external(vi_sim_s* h) {
  if (h->ext_pending) {
    VI_HIP(hipStreamWaitEvent(h->stream, h->ext_ev, 0));
    h->ext_pending = false;
  }
  return 0;
}

int upload_plane(hipStream_t s, void* dst, size_t ld, const void* src, int nx, int nz,
                 size_t esz) {
  VI_HIP(hipMemcpy2DAsync(dst, ld * esz, src, nx * esz, nx * esz, nz, hipMemcpyHostToDevice,
                          s));
  return 0;
}
int download_plane(hipStream_t s, void* dst, const void* src, size_t ld, int nx, int nz,
                   size_t esz) {
  VI_HIP(hipMemcpy2DAsync(dst, nx * esz, src, ld * esz, nx * esz, nz, hipMemcpyDeviceToHost,
                          s));
  return 0;
}

int tile_of(const vi_sim_s* h, int ix, int iz) {
  return ((iz - 4) / (vi::NW * h->rz)) * h->ntx + (ix - 4) / vi::TX;
}

// Per-tile CSR table [off(ntiles+1) | cell(n) | a(n) | original index(n)], uploaded; *order_out
// (optional) receives the original index of each CSR entry.
int upload_table(vi_sim_s* h, DevBuf& buf, int n, const int* ix, const int* iz, const int* a,
                 std::vector<int>* order_out = nullptr) {
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
    cb[k] = q;
  }
  VI_TRY(buf.alloc(tab.size() * sizeof(int), h->stream));
  VI_HIP(hipMemcpyAsync(buf.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice,
                        h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  if (order_out) *order_out = order;
  return 0;
}

// The wavelet samples transposed to [snt][nsrc] with the sources in CSR order, so a tile's
// list loads of one step are contiguous and need no source index first.
int upload_src_wc(vi_sim_s* h, const std::vector<int>& order) {
  const size_t E = h->esz, n = (size_t)h->nsrc, nt = (size_t)h->snt;
  std::vector<char> t(n * nt * E);
  for (size_t k = 0; k < n; ++k)
    for (size_t st = 0; st < nt; ++st)
      std::memcpy(t.data() + (st * n + k) * E, h->src_w_host.data() + ((size_t)order[k] * nt + st) * E, E);
  VI_TRY(h->src_wc.alloc(t.size(), h->stream));
  if (!t.empty())
    VI_HIP(hipMemcpyAsync(h->src_wc.p, t.data(), t.size(), hipMemcpyHostToDevice, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

bool interior(const vi_sim_s* h, int ix, int iz) {
  return ix >= 4 && ix < h->g.nx - 4 && iz >= 4 && iz < h->g.nz - 4;
}

// Re-tile the interior for rows-per-wave rz and rebuild the per-tile tables.
int apply_tiling(vi_sim_s* h, int rz) {
  h->rz = rz;
  h->ntx = (h->g.nx - 8 + vi::TX - 1) / vi::TX;
  h->ntz = (h->g.nz - 8 + vi::NW * rz - 1) / (vi::NW * rz);
  h->ntiles = h->ntx * h->ntz;
  std::vector<int> order;
  VI_TRY(upload_table(h, h->src_tab, h->nsrc, h->src_ix.data(), h->src_iz.data(),
                      h->src_kind.data(), &order));
  VI_TRY(upload_src_wc(h, order));
  VI_TRY(upload_table(h, h->rec_tab, h->nrec, h->rec_ix.data(), h->rec_iz.data(),
                      h->rec_comp.data()));
  touch(h);
  return 0;
}

// Default tiling (measured on MI355X, L = 3 FP32, us per step with 4-, 8- and 16-row tiles;
// the 8- and 16-row tiles run as 8 wavefronts of one and two rows): 256^2 9.3 / 7.9 / 10.1,
// 512^2 11.2 / 10.5 / 11.8, 1024x512 15.1 / 14.6 / 15.8, 1600x600 26.6 / 27.0 / 28.6,
// 1024x2048 51.4 / 52.3 / 56.4, 2048^2 131.0 / 128.1 / 124.7. Counted in 4-row tiles: 8-row
// tiles up to 2048 of them (the launch is latency-bound: fewer, shorter-lived workgroups),
// 16-row tiles from 12288 (halo re-reads amortised), 4-row tiles in between.
int default_rz(int nx, int nz) {
  const long tiles = (long)((nx - 8 + vi::TX - 1) / vi::TX) * ((nz - 8 + vi::NW - 1) / vi::NW);
  return tiles <= 2048 ? 2 : tiles >= 12288 ? 4 : 1;
}

template <typename R>
void build_desc(const vi_sim_s* h, int tile_base, vi::ShotDesc<R>* d) {
  std::memset((void*)d, 0, sizeof(*d));
  d->nx = h->g.nx;
  d->nz = h->g.nz;
  d->ldx = h->ldx;
  d->ldm = h->model->ldm;
  d->ntx = h->ntx;
  d->tile_base = tile_base;
  d->plane = h->plane;
  d->step = h->step;
  d->dt = (R)h->g.dt;
  double c[4];
  vi_fd_weights(h->g.dx, h->g.fd_order_x, c);  // orders validated by vi_create
  for (int l = 0; l < 4; ++l) d->Dx[l] = (R)c[l];
  vi_fd_weights(h->g.dz, h->g.fd_order_z, c);
  for (int l = 0; l < 4; ++l) d->Dz[l] = (R)c[l];
  for (int k = 0; k < h->L; ++k) {
    const R one = R(1), Lr = (R)h->L, tn = (R)h->taun[k];
    d->a1[k] = -one / (Lr * tn);  // visco-kernels.jl:50
    d->a2[k] = -one / tn;         // visco-kernels.jl:51
  }
  d->vx = (R*)h->state[0].p;
  d->vz = (R*)h->state[1].p;
  d->sxx = (R*)h->state[2].p;
  d->szz = (R*)h->state[3].p;
  d->sxz = (R*)h->state[4].p;
  d->M = (R*)h->M.p;
  const ModelStore& m = *h->model;
  d->rho = (const R*)m.arr[0].p + h->model_x0;
  d->lam = (const R*)m.arr[1].p + h->model_x0;
  d->mu = (const R*)m.arr[2].p + h->model_x0;
  d->tau = m.arr[3].p ? (const R*)m.arr[3].p + h->model_x0 : nullptr;
  d->vti = m.vti ? 1 : 0;
  d->c11 = m.vti ? (const R*)m.arr[4].p + h->model_x0 : nullptr;
  d->c33 = m.vti ? (const R*)m.arr[5].p + h->model_x0 : nullptr;
  if (h->cpml_on) {
    d->T = h->cpml.thickness;
    d->xlo = h->cpml.left;
    d->xhi = h->cpml.right;
    d->zlo = h->cpml.top;
    d->zhi = h->cpml.bottom;
    const R* cb = (const R*)h->coef.p;
    const int nx = h->g.nx, nz = h->g.nz;
    d->ax = cb;
    d->bx = cb + nx;
    d->kx = cb + 2 * nx;
    d->az = cb + 3 * nx;
    d->bz = cb + 3 * nx + nz;
    d->kz = cb + 3 * nx + 2 * nz;
    for (int q = 0; q < 8; ++q) d->psi[q] = (R*)h->psi[q].p;
  }
  d->fs = h->fs;
  d->afda = h->fs >= 2 ? (const R*)h->afda.p : nullptr;
  const int* st = (const int*)h->src_tab.p;
  d->src_off = st;
  d->src_cell = st + (h->ntiles + 1);
  d->src_kind = st + (h->ntiles + 1) + h->nsrc;
  d->src_row = st + (h->ntiles + 1) + 2 * h->nsrc;
  d->src_w = (const R*)h->src_w.p;
  d->src_wc = (const R*)h->src_wc.p;
  d->nsrc = h->nsrc;
  d->snt = h->snt;
  const int* rt = (const int*)h->rec_tab.p;
  d->rec_off = rt;
  d->rec_cell = rt + (h->ntiles + 1);
  d->rec_comp = rt + (h->ntiles + 1) + h->nrec;
  d->rec_id = rt + (h->ntiles + 1) + 2 * h->nrec;
  d->traces = (R*)h->traces.p;
  d->nrec = h->nrec;
  d->rec_nt = h->rec_nt;
}

size_t desc_size(size_t esz) {
  return esz == 4 ? sizeof(vi::ShotDesc<float>) : sizeof(vi::ShotDesc<double>);
}

void build_desc_bytes(const vi_sim_s* h, int tile_base, char* out) {
  if (h->esz == 4) build_desc<float>(h, tile_base, (vi::ShotDesc<float>*)out);
  else build_desc<double>(h, tile_base, (vi::ShotDesc<double>*)out);
}

int check_ready(const vi_sim_s* h) {
  if (!h->model) return fail(VI_ESTATE, "vi_step: no model set (vi_set_model)");
  if (h->L && !h->relax_set)
    return fail(VI_ESTATE, "vi_step: vi_set_relaxation has not been called");
  return 0;
}

// Upload the handle's own descriptor and tile->shot table if anything changed.
int refresh_own_desc(vi_sim_s* h) {
  if (!h->desc_dirty) return 0;
  VI_HIP(hipStreamSynchronize(h->stream));
  const size_t ds = desc_size(h->esz);
  h->desc_host.assign(ds, 0);
  build_desc_bytes(h, 0, h->desc_host.data());
  VI_TRY(h->desc_dev.reserve(ds));
  VI_HIP(hipMemcpyAsync(h->desc_dev.p, h->desc_host.data(), ds, hipMemcpyHostToDevice,
                        h->stream));
  const size_t tb = (size_t)h->ntiles * sizeof(int);
  if (h->tile_shot_dev.bytes < tb || !h->tile_shot_dev.p)
    VI_TRY(h->tile_shot_dev.alloc(tb, h->stream));  // zero = shot 0
  VI_HIP(hipStreamSynchronize(h->stream));
  h->desc_dirty = false;
  return 0;
}

template <typename K, typename... Args>
void launch(K kern, dim3 grid, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
            Args... args) {
  if (e0) hipExtLaunchKernelGGL(kern, grid, block, 0, s, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kern, grid, block, 0, s, args...);
}

template <typename R, int RZ, bool AFDA>
int launch_kernel(hipStream_t s, int L, int k, const void* descs, const int* tile_shot,
                  int ntiles, int kstep, hipEvent_t e0, hipEvent_t e1) {
  const dim3 grid(ntiles);
  const vi::ShotDesc<R>* d = (const vi::ShotDesc<R>*)descs;
  const vi::RecSlot rs{};
  // 8- and 16-row tiles (the 2- and 4-rows-per-wave tilings) run as 8 wavefronts of one and
  // two rows: half as much work per wave as 4 wavefronts, so each workgroup finishes sooner
  // (1024x512: 14.5 against 17.8 us per step with 8-row tiles, and 15.1 us with 4-row tiles;
  // 16-row tiles 15.8 against 24.7 us)
  constexpr int KRZ = RZ == 2 ? 1 : RZ == 4 ? 2 : RZ;
  constexpr int KNW = RZ == 2 || RZ == 4 ? 8 : vi::NW;
  const dim3 kblock(vi::TX, KNW);
  if (k == 0) {
    launch(vi::k_velocity<R, KRZ, false, AFDA, KNW>, grid, kblock, s, e0, e1, d, tile_shot,
           kstep, rs);
  } else {
    switch (L) {
#define VI_STRESS(l)                                                                          \
  case l:                                                                                     \
    launch(vi::k_stress<R, l, KRZ, false, AFDA, KNW>, grid, kblock, s, e0, e1, d, tile_shot,  \
           kstep, rs);                                                                        \
    break;
      VI_STRESS(0) VI_STRESS(1) VI_STRESS(2) VI_STRESS(3) VI_STRESS(4) VI_STRESS(5) VI_STRESS(6)
#undef VI_STRESS
      default: return fail(VI_EINVAL, "nrelax out of range");
    }
  }
  VI_HIP(hipGetLastError());
  return 0;
}

// One half step of one shot (null tile table) that also records the forward state of the
// step into a slot (recording variants exist for 1 and 2 rows per wavefront).
template <typename R, int RZ, bool AFDA>
int launch_rec_kernel(hipStream_t s, int L, int k, const void* descs, int ntiles, int kstep,
                      const vi::RecSlot& rs) {
  // 8-row tiles as 8 wavefronts of one row, as in the plain launches
  constexpr int KRZ = RZ == 2 ? 1 : RZ, KNW = RZ == 2 ? 8 : vi::NW;
  const dim3 block(vi::TX, KNW);
  const dim3 grid(ntiles);
  const vi::ShotDesc<R>* d = (const vi::ShotDesc<R>*)descs;
  const int* ts = nullptr;
  if (k == 0) {
    hipLaunchKernelGGL((vi::k_velocity<R, KRZ, true, AFDA, KNW>), grid, block, 0, s, d, ts, kstep,
                       rs);
  } else {
    switch (L) {
#define VI_STRESS(l)                                                                          \
  case l:                                                                                     \
    hipLaunchKernelGGL((vi::k_stress<R, l, KRZ, true, AFDA, KNW>), grid, block, 0, s, d, ts,   \
                       kstep, rs);                                                            \
    break;
      VI_STRESS(0) VI_STRESS(1) VI_STRESS(2) VI_STRESS(3) VI_STRESS(4) VI_STRESS(5) VI_STRESS(6)
#undef VI_STRESS
      default: return fail(VI_EINVAL, "nrelax out of range");
    }
  }
  VI_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int launch_rec(hipStream_t s, int rz, int L, int k, const void* descs, int ntiles, int kstep,
               const vi::RecSlot& rs, bool afda) {
  if (rz == 1)
    return afda ? launch_rec_kernel<R, 1, true>(s, L, k, descs, ntiles, kstep, rs)
                : launch_rec_kernel<R, 1, false>(s, L, k, descs, ntiles, kstep, rs);
  if (rz == 2)
    return afda ? launch_rec_kernel<R, 2, true>(s, L, k, descs, ntiles, kstep, rs)
                : launch_rec_kernel<R, 2, false>(s, L, k, descs, ntiles, kstep, rs);
  return fail(VI_EINVAL, "recording launches need 1 or 2 rows per wavefront");
}

// afda: some shot of the launch has the H-AFDA surface (the variant with those rows)
template <typename R>
int launch_one(hipStream_t s, int rz, int L, int k, const void* descs, const int* tile_shot,
               int ntiles, int kstep, hipEvent_t e0, hipEvent_t e1, bool afda) {
#define VI_RZ(r)                                                                              \
  case r:                                                                                     \
    return afda ? launch_kernel<R, r, true>(s, L, k, descs, tile_shot, ntiles, kstep, e0, e1) \
                : launch_kernel<R, r, false>(s, L, k, descs, tile_shot, ntiles, kstep, e0, e1);
  switch (rz) { VI_RZ(1) VI_RZ(2) VI_RZ(4) VI_RZ(8) }
#undef VI_RZ
  return fail(VI_EINVAL, "bad tiling");
}

// nsteps full steps of one group of shots (same precision and L), then advance their step
// indices on the device.
int launch_steps(hipStream_t s, size_t esz, int rz, int L, void* descs, const int* tile_shot,
                 int ntiles, int nshots, int nsteps, Prof* prof, bool afda) {
  if (nshots == 1) tile_shot = nullptr;  // every tile belongs to shot 0
  for (int q = 0; q < nsteps; ++q) {
    const bool timed = prof && prof->tick();
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (timed) VI_TRY(prof->begin(k, &e0, &e1));
      VI_TRY(esz == 4 ? launch_one<float>(s, rz, L, k, descs, tile_shot, ntiles, q, e0, e1, afda)
                      : launch_one<double>(s, rz, L, k, descs, tile_shot, ntiles, q, e0, e1, afda));
    }
  }
  if (nsteps > 0) {
    const int bs = 64, nb = (nshots + bs - 1) / bs;
    if (esz == 4)
      hipLaunchKernelGGL(vi::k_advance<float>, dim3(nb), dim3(bs), 0, s,
                         (vi::ShotDesc<float>*)descs, nshots, nsteps);
    else
      hipLaunchKernelGGL(vi::k_advance<double>, dim3(nb), dim3(bs), 0, s,
                         (vi::ShotDesc<double>*)descs, nshots, nsteps);
    VI_HIP(hipGetLastError());
  }
  return 0;
}

int step_single(vi_sim_s* h, int nsteps) {
  VI_TRY(check_ready(h));
  VI_TRY(wait_external(h));
  VI_TRY(refresh_own_desc(h));
  VI_TRY(launch_steps(h->stream, h->esz, h->rz, h->L, h->desc_dev.p,
                      (const int*)h->tile_shot_dev.p, h->ntiles, 1, nsteps, &h->prof,
                      h->fs >= 2));
  h->step += nsteps;
  h->own_pending = true;
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

std::vector<DevBuf*> mutable_arrays(vi_sim_s* h) {
  std::vector<DevBuf*> v;
  for (auto& b : h->state) v.push_back(&b);
  v.push_back(&h->M);
  for (auto& b : h->psi) v.push_back(&b);
  v.push_back(&h->traces);
  return v;
}

// src = {rho, lam | C13, mu | C55, tau, C11, C33}; C11 and C33 given only for VTI
int make_model_store(int nx, int nz, size_t esz, bool with_tau, const void* const src[6],
                     hipStream_t s, std::shared_ptr<ModelStore>* out) {
  auto m = std::make_shared<ModelStore>();
  m->nx = nx;
  m->nz = nz;
  m->esz = esz;
  m->ldm = (int)round_up((size_t)nx + 72, 64);
  m->nzp = nz + vi::NW * vi::RZ_MAX + 16;
  m->has_tau = with_tau;
  m->vti = src[4] != nullptr;
  const size_t pb = (size_t)m->ldm * m->nzp * esz;
  if (pb >= (size_t(1) << 31)) return fail(VI_EINVAL, "model plane must stay below 2 GiB");
  for (int q = 0; q < 6; ++q) {
    if (q == 3 && !with_tau) continue;
    if (q >= 4 && !m->vti) continue;
    VI_TRY(m->arr[q].alloc(pb, s));
    VI_TRY(upload_plane(s, m->arr[q].p, m->ldm, src[q], nx, nz, esz));
  }
  VI_HIP(hipStreamSynchronize(s));
  *out = m;
  return 0;
}

// ---------------------------------------------------------------- batches
void invalidate_graph(vi_batch_s* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
  b->graph = nullptr;
  b->graph_def = nullptr;
}

// Bytes one step of a shot streams through: its state, memory variables, CPML memory and
// the columns of the model it reads.
size_t working_set_bytes(vi_sim_s* h) {
  size_t t = 0;
  for (auto* b : mutable_arrays(h))
    if (b != &h->traces) t += b->bytes;
  const ModelStore& m = *h->model;
  for (auto& a : m.arr)
    if (a.p) t += (size_t)m.nzp * (size_t)(h->g.nx + 8) * m.esz;
  return t;
}

int batch_build(vi_batch_s* b) {
  invalidate_graph(b);
  b->groups.clear();
  // one launch group per (precision, L) class; with cache blocking, each class is packed into
  // groups of at most block_bytes of working set (at least one shot each), and every group
  // runs all steps of a block before the next starts, so a group's data stays in the
  // Infinity Cache across its launches (shots are independent, nothing else changes). The
  // packing is first fit by decreasing working set: fewer, fuller groups than cutting the
  // member list in order, so fewer launches per step.
  std::map<std::pair<int, int>, std::vector<vi_sim_s*>> classes;
  for (auto* h : b->members) {
    VI_TRY(check_ready(h));
    if (h->rz != b->rz) VI_TRY(apply_tiling(h, b->rz));
    classes[std::make_pair((int)h->esz, h->L)].push_back(h);
  }
  // n groups run at once on n streams: each gets 1/n of the block
  const size_t limit = b->block_bytes / (size_t)b->nstreams;
  for (auto& cl : classes) {
    std::vector<std::pair<size_t, vi_sim_s*>> order;  // (working set, shot), largest first
    for (auto* h : cl.second) order.emplace_back(working_set_bytes(h), h);
    if (limit)
      std::stable_sort(order.begin(), order.end(),
                       [](const auto& x, const auto& y) { return x.first > y.first; });
    const size_t first = b->groups.size();
    std::vector<size_t> filled;
    for (const auto& o : order) {
      size_t q = 0;
      while (q < filled.size() && limit && filled[q] + o.first > limit) ++q;
      if (q == filled.size()) {
        b->groups.emplace_back(new BatchGroup());
        b->groups.back()->esz = cl.first.first;
        b->groups.back()->L = cl.first.second;
        filled.push_back(0);
      }
      b->groups[first + q]->members.push_back(o.second);
      filled[q] += o.first;
    }
  }
  for (auto& gp : b->groups) {
    BatchGroup& g = *gp;
    const size_t ds = desc_size(g.esz);
    g.desc_host.assign(ds * g.members.size(), 0);
    std::vector<int> ts;
    int base = 0;
    for (size_t q = 0; q < g.members.size(); ++q) {
      build_desc_bytes(g.members[q], base, g.desc_host.data() + q * ds);
      ts.insert(ts.end(), g.members[q]->ntiles, (int)q);
      base += g.members[q]->ntiles;
    }
    g.ntiles = base;
    g.afda = false;
    for (auto* m : g.members) g.afda = g.afda || m->fs >= 2;
    VI_TRY(g.desc_dev.reserve(g.desc_host.size()));
    VI_TRY(g.tile_shot_dev.reserve(ts.size() * sizeof(int)));
    VI_HIP(hipMemcpyAsync(g.desc_dev.p, g.desc_host.data(), g.desc_host.size(),
                          hipMemcpyHostToDevice, b->stream));
    VI_HIP(hipMemcpyAsync(g.tile_shot_dev.p, ts.data(), ts.size() * sizeof(int),
                          hipMemcpyHostToDevice, b->stream));
    VI_HIP(hipStreamSynchronize(b->stream));
  }
  b->versions.resize(b->members.size());
  b->steps.resize(b->members.size());
  for (size_t q = 0; q < b->members.size(); ++q) {
    b->versions[q] = b->members[q]->version;
    b->steps[q] = b->members[q]->step;
  }
  b->built = true;
  return 0;
}

int batch_launch(vi_batch_s* b, int nsteps) {
  const int ns = (int)std::min<size_t>((size_t)b->nstreams, b->groups.size());
  if (ns > 1) {  // fork: the side streams start after everything queued on the batch stream
    VI_HIP(hipEventRecord(b->fork_ev, b->stream));
    for (int k = 0; k < ns - 1; ++k) VI_HIP(hipStreamWaitEvent(b->side[k], b->fork_ev, 0));
  }
  int rc = 0;
  for (size_t q = 0; q < b->groups.size() && !rc; ++q) {
    BatchGroup& g = *b->groups[q];
    const int k = (int)(q % (size_t)ns);
    rc = launch_steps(k ? b->side[k - 1] : b->stream, g.esz, b->rz, g.L, g.desc_dev.p,
                      (const int*)g.tile_shot_dev.p, g.ntiles, (int)g.members.size(), nsteps,
                      &b->prof, g.afda);
  }
  for (int k = 0; k < ns - 1; ++k) {  // join, also after an error: no work is left behind
    VI_HIP(hipEventRecord(b->join_ev[k], b->side[k]));
    VI_HIP(hipStreamWaitEvent(b->stream, b->join_ev[k], 0));
  }
  return rc;
}

int batch_prepare(vi_batch_s* b) {
  for (auto* h : b->members)
    if (h->grad_busy) return fail(VI_ESTATE, "batch member in use by a gradient job");
  bool stale = !b->built;
  // a member stepped on its own since the last batch step has another step index than the
  // batch descriptors (sources and receivers would use the wrong time sample)
  for (size_t q = 0; !stale && q < b->members.size(); ++q)
    stale = b->versions[q] != b->members[q]->version || b->steps[q] != b->members[q]->step;
  if (stale) {
    for (auto* h : b->members) {
      VI_TRY(wait_external(h));
      VI_HIP(hipStreamSynchronize(h->stream));
      h->own_pending = false;
    }
    VI_HIP(hipStreamSynchronize(b->stream));
    VI_TRY(batch_build(b));
  }
  // order after work queued on the members' own streams
  for (auto* h : b->members) {
    if (h->own_pending) {
      VI_HIP(hipEventRecord(h->own_ev, h->stream));
      VI_HIP(hipStreamWaitEvent(b->stream, h->own_ev, 0));
      h->own_pending = false;
    }
  }
  return 0;
}

int batch_finish(vi_batch_s* b, int nsteps) {
  for (size_t q = 0; q < b->members.size(); ++q) {
    vi_sim_s* h = b->members[q];
    h->step += nsteps;
    b->steps[q] = h->step;
    h->desc_dirty = true;  // its own descriptor holds the old step
    VI_HIP(hipEventRecord(h->ext_ev, b->stream));
    h->ext_pending = true;
  }
  return 0;
}


// ---------------------------------------------------------------- rates (right-hand side)
// Workspace of one vi_rates call: the 5 + 3 L rate planes (state pitch) and the source list
// (ix, iz, kind and value per source).
size_t rates_bytes(const vi_sim_s* h) {
  const size_t planes = (size_t)(5 + 3 * h->L) * h->plane * h->esz;
  return planes + (size_t)h->nsrc * (3 * sizeof(int) + h->esz) + 256;
}

template <typename R, int L>
void launch_rates_l(const vi::ShotDesc<R>& d, const int* six, const int* siz, const int* sk,
                    const R* sv, int n, R* out, hipStream_t st) {
  const dim3 grid((d.nx + 63) / 64, (d.nz + 3) / 4), block(64, 4);
  hipLaunchKernelGGL((vi::k_rates<R, L>), grid, block, 0, st, d, six, siz, sk, sv, n, out);
}

// The rates of the next half steps at the handle's state into the host arrays dv [2][nz][nx],
// ds [3][nz][nx] and dm [3][L][nz][nx] (optional), with ws (>= rates_bytes, 256-byte
// aligned) as the device workspace. Allocates nothing: the source list is copied from the
// handle's own host arrays and the caller's values.
template <typename R>
int rates_run(vi_sim_s* h, const void* src_values, void* dv, void* ds, void* dm, char* ws) {
  hipStream_t st = h->stream;
  vi::ShotDesc<R> d;
  build_desc<R>(h, 0, &d);
  const size_t E = sizeof(R), P = h->plane;
  R* out = (R*)ws;
  const size_t planes_b = (size_t)(5 + 3 * h->L) * P * E;
  const int n = h->nsrc;
  int* six = (int*)(ws + planes_b);
  int* siz = six + n;
  int* sk = siz + n;
  R* sv = (R*)(ws + ((planes_b + 3 * (size_t)n * sizeof(int) + 63) & ~(size_t)63));
  if (n) {
    VI_HIP(hipMemcpyAsync(six, h->src_ix.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
    VI_HIP(hipMemcpyAsync(siz, h->src_iz.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
    VI_HIP(hipMemcpyAsync(sk, h->src_kind.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
    if (src_values) VI_HIP(hipMemcpyAsync(sv, src_values, n * E, hipMemcpyHostToDevice, st));
  }
  const R* svp = src_values ? sv : nullptr;
  switch (h->L) {
#define VI_RL(l) case l: launch_rates_l<R, l>(d, six, siz, sk, svp, n, out, st); break;
    VI_RL(0) VI_RL(1) VI_RL(2) VI_RL(3) VI_RL(4) VI_RL(5) VI_RL(6)
#undef VI_RL
    default: return fail(VI_EINVAL, "nrelax out of range");
  }
  VI_HIP(hipGetLastError());
  if (h->fs == 1) {
    hipLaunchKernelGGL(vi::k_rates_images<R>, dim3((h->g.nx - 8 + 63) / 64), dim3(64, 4), 0, st,
                       h->g.nx, h->ldx, P, out);
    VI_HIP(hipGetLastError());
  }
  const int nx = h->g.nx, nz = h->g.nz;
  const size_t nc = (size_t)nx * nz;
  for (int q = 0; q < 2; ++q)
    VI_TRY(download_plane(st, (char*)dv + q * nc * E, out + q * P, h->ldx, nx, nz, E));
  for (int q = 0; q < 3; ++q)
    VI_TRY(download_plane(st, (char*)ds + q * nc * E, out + (2 + q) * P, h->ldx, nx, nz, E));
  if (dm)
    for (int q = 0; q < 3 * h->L; ++q)
      VI_TRY(download_plane(st, (char*)dm + q * nc * E, out + (5 + q) * P, h->ldx, nx, nz, E));
  VI_HIP(hipStreamSynchronize(st));
  h->own_pending = false;
  return 0;
}

// Device workspace of vi_rates_noalloc: a fixed array in the library's code object, so the
// entry point never allocates. Each GPU that loads the library holds its own copy until the
// process ends; host code must use its per-device address from hipGetSymbolAddress, never
// the array's name. Calls that overlap on one GPU share it and race.
constexpr size_t kRatesScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) char scratch[kRatesScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// The address of `scratch` on `device`, looked up once per device with it current.
int scratch_on(int device, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return fail(VI_EDEVICE, "device index out of range");
  if (!g_scratch[device]) {
    int cur = 0;
    VI_HIP(hipGetDevice(&cur));
    if (cur != device) VI_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(VI_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (char*)p;
  }
  *out = g_scratch[device];
  return 0;
}

int rates_checks(vi_sim_s* h, const void* dv, const void* ds, const char* who) {
  VI_TRY(check_handle(h));
  if (!dv || !ds) return fail(VI_EINVAL, std::string(who) + ": dv and ds are required");
  VI_TRY(check_ready(h));
  VI_TRY(wait_external(h));
  return 0;
}

// ---------------------------------------------------------------- gradient (adjoint)
// Device buffers of one gradient call: checkpoints (forward state every K steps), the
// segment slots (per step what the adjoint needs from the forward pass: the stress divergence
// fx, fz, the strain rates exx, ezz, gxz and the memory-variable sums; see vi::RecSlot), the
// adjoint state with work planes and accumulators, the residual and the injection lists.
// Regions inside the checkpoint buffer start on 256-byte boundaries.
struct GradLayout {
  int K = 1, nck = 1;
  size_t state_bytes = 0, slot_bytes = 0, psi_total = 0, guard = 0, adj_bytes = 0;
  size_t resid_bytes = 0, lists_bytes = 0;
  size_t total() const {
    return state_bytes * nck + slot_bytes * K + adj_bytes + resid_bytes + lists_bytes;
  }
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

void swap_buf(DevBuf& a, DevBuf& b) {
  std::swap(a.p, b.p);
  std::swap(a.bytes, b.bytes);
  std::swap(a.cap, b.cap);
  std::swap(a.device, b.device);
}

// live[] = byte sizes of the checkpointed arrays (5 state planes, M, 8 psi strips), in the
// order the gradient copies them; lists_ints bounds the injection lists
GradLayout grad_layout(int nt, int K, int L, size_t pb, size_t ldx, size_t esz,
                       const size_t live[14], int nrec, size_t lists_ints) {
  GradLayout g;
  if (K <= 0) K = std::max(1, (int)std::ceil(std::sqrt((double)nt)));
  g.K = std::min(K, nt);
  g.nck = (nt + g.K - 1) / g.K;
  for (int q = 0; q < 14; ++q) g.state_bytes += align256(live[q]);
  g.slot_bytes = (L ? 8 : 5) * pb;
  for (int q = 0; q < 8; ++q) g.psi_total += live[6 + q];
  g.guard = 4 * ldx * esz;
  const size_t nplanes_adj = 5 + 3 * (size_t)L + 4 + 4 + 7 + 2;
  g.adj_bytes = nplanes_adj * pb + 8 * g.guard + g.psi_total;
  g.resid_bytes = 2 * align256((size_t)nt * nrec * esz) + ((size_t)nrec + 1) * sizeof(double);
  g.lists_bytes = lists_ints * sizeof(int);
  return g;
}

// A gradient computation as a sequence of asynchronous chunks on the handle's stream:
//   forward pass storing a checkpoint every K steps -> residual and misfit on the device ->
//   for each segment, last first: restore its checkpoint, recompute it storing every step's
//   forward quantities in the slots, then the transposed steps backwards -> finalisation of
//   the accumulators.
// The last segment is recorded into the slots while the forward pass computes it, so it is
// never recomputed; with one segment (K >= nt) nothing is recomputed at all.
// advance() enqueues at most max_steps forward, recomputed or transposed steps and never
// waits for the device, so one host thread can interleave many gradients with other work on
// the GPU.
template <typename R>
struct GradRun : vi_grad_s {
  int nt = 0, K = 1, nck = 1, L = 0, nrec = 0, nsrc = 0, ntgt = 0, ntgt_v = 0;
  size_t E = sizeof(R), pb = 0;
  GradLayout lay;
  size_t psi_bytes[8] = {};
  std::vector<DevBuf*> live;
  DevBuf ck, seg, adj, resid, lists;
  double J = 0.0;
  R* dobs_dev = nullptr;    // in resid, after the residual: the observed data [nrec][nt]
  double* Jdev = nullptr;   // in resid, after the observed data: J, then nrec partial sums
  vi::AdjDesc<R> d{};
  R* g_rho_out = nullptr;
  R* g_mu_out = nullptr;
  const int *t_cell = nullptr, *t_comp = nullptr, *t_ptr = nullptr, *t_rec = nullptr;
  enum Phase { FWD, RESTORE, RECOMP, BACK, DONE } ph = FWD;
  int rz0 = 0;  // the handle's tiling before the job (restored at the end)
  int s = 0, n = 0;  // segment, step (FWD: steps done in segment s)

  ~GradRun() override {
    if (h) {  // the workspace goes back to the handle for its next gradient
      (void)hipStreamSynchronize(h->stream);
      swap_buf(ck, h->grad_ck);
      swap_buf(seg, h->grad_seg);
      swap_buf(adj, h->grad_adj);
    }
  }

  struct Cp { const void* s; void* d; size_t b; };
  int run_copies(const std::vector<Cp>& v) {
    for (size_t i = 0; i < v.size(); i += 16) {
      vi::CopyList c{};
      size_t mx = 0;
      c.n = (int)std::min<size_t>(16, v.size() - i);
      for (int k = 0; k < c.n; ++k) {
        c.src[k] = (const char*)v[i + k].s;
        c.dst[k] = (char*)v[i + k].d;
        c.bytes[k] = v[i + k].b;
        mx = std::max(mx, v[i + k].b);
      }
      const unsigned gx = (unsigned)std::min<size_t>(1024, std::max<size_t>(1, (mx / 16 + 255) / 256));
      hipLaunchKernelGGL(vi::k_copy_list, dim3(gx, c.n), dim3(256), 0, h->stream, c);
    }
    VI_HIP(hipGetLastError());
    return 0;
  }
  int copy_state(int seg_idx, bool to_ck) {
    char* dst = (char*)ck.p + (size_t)seg_idx * lay.state_bytes;
    std::vector<Cp> v;
    size_t off = 0;
    for (auto* b : live) {
      if (!b->p) continue;
      if (to_ck) v.push_back({b->p, dst + off, b->bytes});
      else v.push_back({dst + off, b->p, b->bytes});
      off += align256(b->bytes);
    }
    return run_copies(v);
  }

  int init(vi_sim_s* hh, int nt_, const R* dobs_, int K_) {
    h = hh;
    nt = nt_;
    L = h->L;
    pb = h->plane * E;
    nrec = h->nrec;
    nsrc = h->nsrc;
    hipStream_t st = h->stream;
    VI_TRY(vi_reset(h));
    rz0 = h->rz;
    if (h->rz > 2) VI_TRY(apply_tiling(h, 2));  // the recording kernels use short tiles
    VI_TRY(refresh_own_desc(h));
    live = mutable_arrays(h);
    live.pop_back();  // traces are not part of the checkpointed state
    size_t live_bytes[14];
    for (int q = 0; q < 14; ++q) live_bytes[q] = live[q]->bytes;
    for (int q = 0; q < 8; ++q) psi_bytes[q] = live_bytes[6 + q];
    // injection targets: distinct (component, cell) pairs with their receivers in ascending
    // order; layout [tcell | tcomp | tptr (ntgt+1) | trec (nrec)]
    std::map<std::pair<int, int>, std::vector<int>> targets;
    for (int r = 0; r < nrec; ++r)
      targets[{h->rec_comp[r], h->rec_iz[r] * h->ldx + h->rec_ix[r]}].push_back(r);
    ntgt = (int)targets.size();
    ntgt_v = 0;  // targets of the velocity components come first (the map orders by component)
    for (auto& kv : targets) ntgt_v += kv.first.first <= 1;
    std::vector<int> lst(3 * (size_t)ntgt + 1 + (size_t)nrec);
    lay = grad_layout(nt, K_, L, pb, h->ldx, E, live_bytes, nrec, lst.size());
    K = lay.K;
    nck = lay.nck;
    // the handle's workspace from an earlier gradient, grown when too small; checkpoints
    // and slots are written before they are read, so only the adjoint part is cleared
    swap_buf(ck, h->grad_ck);
    swap_buf(seg, h->grad_seg);
    swap_buf(adj, h->grad_adj);
    VI_TRY(ck.reserve(lay.state_bytes * nck));
    VI_TRY(seg.reserve(lay.slot_bytes * K));
    // adjoint state, work arrays and accumulators; ehat/dhat planes get 4 zero guard rows in
    // front: the transposed z-operators of the free-surface image rows (rows 0..3) read rows
    // -4..-1, where the derivative adjoints are 0
    VI_TRY(adj.reserve(lay.adj_bytes));
    VI_HIP(hipMemsetAsync(adj.p, 0, lay.adj_bytes, st));
    char* ab = (char*)adj.p;
    auto take = [&](size_t bytes) { char* p = ab; ab += bytes; return (R*)p; };
    auto take_guarded = [&]() { ab += lay.guard; return take(pb); };
    d.avx = take(pb); d.avz = take(pb); d.asxx = take(pb); d.aszz = take(pb); d.asxz = take(pb);
    d.aM = take(3 * (size_t)L * pb);
    for (int q = 0; q < 4; ++q) d.ehat[q] = take_guarded();
    for (int q = 0; q < 4; ++q) d.dhat[q] = take_guarded();
    d.g_lam = take(pb); d.g_mu = take(pb); d.g_muxz = take(pb); d.g_tau = take(pb);
    d.g_bx = take(pb); d.g_bz = take(pb); d.g_c33 = take(pb);
    g_rho_out = take(pb);
    g_mu_out = take(pb);
    for (int q = 0; q < 8; ++q) d.apsi[q] = take(psi_bytes[q]);
    // [residual [nt][nrec] | observed data [nrec][nt] | J, nrec partial sums]
    VI_TRY(resid.reserve(lay.resid_bytes));
    dobs_dev = (R*)((char*)resid.p + align256((size_t)nt * nrec * E));
    VI_HIP(hipMemcpyAsync(dobs_dev, dobs_, (size_t)nt * nrec * E, hipMemcpyHostToDevice, st));
    // the caller may release dobs as soon as begin returns (a chunked job keeps no host
    // pointer), so the copy completes here
    VI_HIP(hipStreamSynchronize(st));
    Jdev = (double*)((char*)dobs_dev + align256((size_t)nt * nrec * E));
    {
      int t = 0, k = 0;
      int* tptr = lst.data() + 2 * ntgt;
      int* trec = tptr + ntgt + 1;
      for (auto& kv : targets) {
        lst[t] = kv.first.second;
        lst[ntgt + t] = kv.first.first;
        tptr[t] = k;
        for (int r : kv.second) trec[k++] = r;
        ++t;
      }
      tptr[ntgt] = k;
    }
    VI_TRY(lists.alloc(lst.size() * sizeof(int), st));
    VI_HIP(hipMemcpyAsync(lists.p, lst.data(), lst.size() * sizeof(int), hipMemcpyHostToDevice, st));
    t_cell = (const int*)lists.p;
    t_comp = t_cell + ntgt;
    t_ptr = t_comp + ntgt;
    t_rec = t_ptr + ntgt + 1;
    d.tcell = t_cell;
    d.tcomp = t_comp;
    d.tptr = t_ptr;
    d.trec = t_rec;
    d.ntgt_v = ntgt_v;
    // static part of the adjoint descriptor
    d.nx = h->g.nx; d.nz = h->g.nz; d.ldx = h->ldx; d.ldm = h->model->ldm;
    {
      const vi::ShotDesc<R>* sd = (const vi::ShotDesc<R>*)h->desc_host.data();
      d.dt = sd->dt;
      for (int l = 0; l < 4; ++l) { d.Dx[l] = sd->Dx[l]; d.Dz[l] = sd->Dz[l]; }
      for (int k = 0; k < L; ++k) { d.a1[k] = sd->a1[k]; d.a2[k] = sd->a2[k]; }
      d.T = sd->T; d.xlo = sd->xlo; d.xhi = sd->xhi; d.zlo = sd->zlo; d.zhi = sd->zhi; d.fs = sd->fs;
      d.afda = sd->afda;
      d.rho = sd->rho; d.lam = sd->lam; d.mu = sd->mu; d.tau = sd->tau;
      d.c11 = sd->c11; d.c33 = sd->c33; d.vti = sd->vti;
      d.ax = sd->ax; d.bx = sd->bx; d.kx = sd->kx; d.az = sd->az; d.bz = sd->bz; d.kz = sd->kz;
    }
    d.L = L;
    d.plane = h->plane;
    VI_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    {  // 2-row tiles only when their launch holds at least four rounds of resident
       // workgroups (3 per CU at the 2-row register count); shorter launches lose more
       // to the last round's tail than the taller tiles save in halo rows (1600x600: 1875
       // 2-row workgroups, 2 % slower per gradient than 1-row tiles; 2048x1024: equal)
      const long ntx = (d.nx - 8 + vi::TX - 1) / vi::TX;
      const long tiles2 = ntx * ((d.nz - 8 + 2 * vi::NW - 1) / (2 * vi::NW));
      arz = tiles2 >= 4L * 3 * 256 ? 2 : 1;
    }
    VI_HIP(hipStreamSynchronize(st));
    h->own_pending = true;
    h->grad_busy = true;
    return 0;
  }

  int arz = 2;  // rows per wavefront of the transposed-step tiles

  template <int RZ>
  int backward_kernels() {
    hipStream_t st = h->stream;
    constexpr int TZ = vi::NW * RZ;
    const dim3 blk(vi::TX, vi::NW);
    const int ntx = (d.nx - 8 + vi::TX - 1) / vi::TX;
    const dim3 gint(ntx * ((d.nz - 8 + TZ - 1) / TZ));
    const int row0 = d.fs == 1 ? 0 : 4;  // image rows only with imaging
    const dim3 gsg(ntx * ((d.nz - 4 - row0 + TZ - 1) / TZ));
    const bool herm = d.fs == 3;  // formula #3 has kernels of its own
#define VI_ADJ_STRESS(l)                                                                    \
  case l:                                                                                   \
    if (herm) hipLaunchKernelGGL((vi::k_adj_stress<R, l, RZ, true>), gint, blk, 0, st, d);  \
    else hipLaunchKernelGGL((vi::k_adj_stress<R, l, RZ, false>), gint, blk, 0, st, d);      \
    break;
    switch (L) {
      VI_ADJ_STRESS(0) VI_ADJ_STRESS(1) VI_ADJ_STRESS(2) VI_ADJ_STRESS(3) VI_ADJ_STRESS(4)
      VI_ADJ_STRESS(5) VI_ADJ_STRESS(6)
    }
#undef VI_ADJ_STRESS
    hipLaunchKernelGGL((vi::k_adj_vel<R, RZ>), gint, blk, 0, st, d);
    hipLaunchKernelGGL((vi::k_adj_sgather<R, RZ>), gsg, blk, 0, st, d, row0);
    VI_HIP(hipGetLastError());
    return 0;
  }

  int backward_step(int step, const char* slot) {
    hipStream_t st = h->stream;
    d.ffx = (const R*)(slot);
    d.ffz = (const R*)(slot + pb);
    d.fexx = (const R*)(slot + 2 * pb);
    d.fezz = (const R*)(slot + 3 * pb);
    d.fgxz = (const R*)(slot + 4 * pb);
    d.fM = (const R*)(slot + 5 * pb);
    d.resid_row = (const R*)resid.p + (size_t)step * nrec;
    // stress-component receivers: their own launch before k_adj_stress reads s_hat (the
    // velocity-component ones are injected by k_adj_stress itself)
    const int ns = ntgt - ntgt_v;
    if (ns > 0)
      hipLaunchKernelGGL(vi::k_adj_inject<R>, dim3((ns + 255) / 256), dim3(256), 0, st, d.avx,
                         d.avz, d.asxx, d.aszz, d.asxz, t_cell + ntgt_v, t_comp + ntgt_v,
                         t_ptr + ntgt_v, t_rec, ns, d.resid_row);
    return arz == 1 ? backward_kernels<1>() : backward_kernels<2>();
  }

  vi::RecSlot slot_of(int k) const {
    vi::RecSlot rs{};
    rs.base = (char*)seg.p + (size_t)k * lay.slot_bytes;
    rs.pb = pb;
    return rs;
  }

  int advance(int max_steps) override {
    hipStream_t st = h->stream;
    int budget = max_steps;
    while (budget > 0 && ph != DONE) {
      if (ph == FWD) {
        const int len = std::min(K, nt - s * K);
        if (s == nck - 1) {
          // last segment: record each step into its slot as it is computed (the descriptor
          // step is s K here, advanced by the plain launches of the earlier segments)
          for (; n < len && budget > 0; ++n, --budget) {
            const vi::RecSlot rs = slot_of(n);
            VI_TRY(launch_rec<R>(st, h->rz, L, 0, h->desc_dev.p, h->ntiles, n, rs, h->fs >= 2));
            VI_TRY(launch_rec<R>(st, h->rz, L, 1, h->desc_dev.p, h->ntiles, n, rs, h->fs >= 2));
          }
        } else {
          if (n == 0) VI_TRY(copy_state(s, true));
          const int m = std::min(len - n, budget);
          VI_TRY(launch_steps(st, E, h->rz, L, h->desc_dev.p, (const int*)h->tile_shot_dev.p,
                              h->ntiles, 1, m, nullptr, h->fs >= 2));
          n += m;
          budget -= m;
        }
        if (n == len) {
          ++s;
          n = 0;
        }
        if (s == nck) {  // residual and misfit on the device, then the transposed steps
          const size_t nn = (size_t)nt * nrec;
          const unsigned gb = (unsigned)std::min<size_t>(4096, (nn + 255) / 256);
          hipLaunchKernelGGL(vi::k_residual<R>, dim3(gb), dim3(256), 0, st,
                             (const R*)h->traces.p, (const R*)dobs_dev, (R*)resid.p, nt, nrec);
          hipLaunchKernelGGL(vi::k_misfit_rec<R>, dim3(nrec), dim3(256), 0, st,
                             (const R*)resid.p, nt, nrec, Jdev + 1);
          hipLaunchKernelGGL(vi::k_misfit_sum, dim3(1), dim3(64), 0, st, (const double*)(Jdev + 1),
                             nrec, Jdev);
          VI_HIP(hipGetLastError());
          s = nck - 1;  // its slots were filled by the forward pass
          n = nt - 1;
          ph = BACK;
        }
      } else if (ph == RESTORE) {
        VI_TRY(copy_state(s, false));
        hipLaunchKernelGGL(vi::k_set_step<R>, dim3(1), dim3(64), 0, st,
                           (vi::ShotDesc<R>*)h->desc_dev.p, 0, (long)s * K);
        VI_HIP(hipGetLastError());
        n = s * K;
        ph = RECOMP;
      } else if (ph == RECOMP) {  // recompute the segment, storing each step's inputs
        const int n0 = s * K, n1 = std::min(n0 + K, nt);
        for (; n < n1 && budget > 0; ++n, --budget) {
          // both half steps record the step's forward state into its slot as they go
          const vi::RecSlot rs = slot_of(n - n0);
          VI_TRY(launch_rec<R>(st, h->rz, L, 0, h->desc_dev.p, h->ntiles, n - n0, rs, h->fs >= 2));
          VI_TRY(launch_rec<R>(st, h->rz, L, 1, h->desc_dev.p, h->ntiles, n - n0, rs, h->fs >= 2));
        }
        if (n == n1) {
          n = n1 - 1;
          ph = BACK;
        }
      } else if (ph == BACK) {
        const int n0 = s * K;
        for (; n >= n0 && budget > 0; --n, --budget)
          VI_TRY(backward_step(n, (const char*)seg.p + (size_t)(n - n0) * lay.slot_bytes));
        if (n < n0) {
          if (s == 0) {
            const dim3 gall((d.nx + 63) / 64, (d.nz + 3) / 4);
            hipLaunchKernelGGL(vi::k_grad_finalize<R>, gall, dim3(64, 4), 0, st, d, g_rho_out,
                               g_mu_out);
            VI_HIP(hipGetLastError());
            ph = DONE;
          } else {
            --s;
            ph = RESTORE;
          }
        }
      }
    }
    done = ph == DONE;
    VI_HIP(hipEventRecord(ev, st));
    return 0;
  }

  int finish(double* misfit, void* const out[6]) override {
    while (ph != DONE) VI_TRY(advance(1 << 30));  // drive the rest synchronously
    hipStream_t st = h->stream;
    const int nx = h->g.nx, nz = h->g.nz;
    const R* planes[6] = {g_rho_out, d.g_lam, g_mu_out, L ? d.g_tau : nullptr,
                          d.vti ? d.g_mu : nullptr, d.vti ? d.g_c33 : nullptr};
    // the gradients and J go through the handle's pinned staging buffer (one strided copy per
    // plane at the full link rate; strided copies into pageable memory go row by row), then
    // to the caller's arrays
    const size_t pe = (size_t)nx * nz * E;
    VI_TRY(h->grad_host.reserve(6 * pe + sizeof(double)));
    char* hs = (char*)h->grad_host.p;
    for (int q = 0; q < 6; ++q)
      if (out && out[q] && planes[q])
        VI_HIP(hipMemcpy2DAsync(hs + q * pe, nx * E, planes[q], h->ldx * E, nx * E, nz,
                                hipMemcpyDeviceToHost, st));
    VI_HIP(hipMemcpyAsync(hs + 6 * pe, Jdev, sizeof(double), hipMemcpyDeviceToHost, st));
    VI_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < 6; ++q)
      if (out && out[q] && planes[q]) std::memcpy(out[q], hs + q * pe, pe);
    std::memcpy(&J, hs + 6 * pe, sizeof(double));
    if (misfit) *misfit = J;
    // leave the handle at a zero state, step 0, with its own tiling
    h->grad_busy = false;
    if (rz0 && h->rz != rz0) VI_TRY(apply_tiling(h, rz0));
    return vi_reset(h);
  }
};

int gradient_begin(vi_sim_s* h, int nt, const void* dobs, int K, vi_grad* out) {
  std::unique_ptr<vi_grad_s> g;
  int rc;
  if (h->esz == 4) {
    auto* r = new GradRun<float>();
    g.reset(r);
    rc = r->init(h, nt, (const float*)dobs, K);
  } else {
    auto* r = new GradRun<double>();
    g.reset(r);
    rc = r->init(h, nt, (const double*)dobs, K);
  }
  if (rc) {
    (void)hipStreamSynchronize(h->stream);
    return rc;
  }
  *out = g.release();
  return 0;
}

}  // namespace

extern "C" {

const char* vi_last_error(void) { return g_err.c_str(); }
const char* vi_version(void) { return "vi_mi355x 0.2 (gfx950)"; }

void vi_fd_coeffs(double h, double out[4]) { vi_fd_weights(h, 8, out); }

// Staggered first-derivative weights of order 2K (K <= 4): _stencil(x, 0, 1)
// (Stencil.jl:9-14) at the nodes x = +-1/2 .. +-(2K-1)/2, as Stencil(order; stagger)
// builds them (Stencil.jl:280-320), folded to the half stencil of diff-funcs.jl:3-32 (the
// weight at -x is minus the weight at +x). Each weight is the derivative of a Lagrange
// basis polynomial at 0, w_m = sum_{q != m} prod_{l != m, q} (0 - x_l) / prod_{l != m}
// (x_m - x_l), in integers of h/2 (every factor at most 14: exact in int64 and in double),
// rounded once by the division and then divided by h. Order 8 gives 1225/1024, -245/3072,
// 49/5120, -5/7168. Unused entries are 0.
int vi_fd_weights(double h, int order, double out[4]) {
  if (!out || !(h > 0) || order < 2 || order > 8 || order % 2)
    return fail(VI_EINVAL, "vi_fd_weights: need h > 0, order 2, 4, 6 or 8 and an output");
  const int K = order / 2, n = order;
  long long X[8];
  for (int l = 0; l < K; ++l) {
    X[l] = 2 * l + 1;       // +(2l+1)/2
    X[K + l] = -(2 * l + 1);  // -(2l+1)/2
  }
  for (int l = 0; l < 4; ++l) out[l] = 0.0;
  for (int m = 0; m < K; ++m) {
    long long num = 0, den = 1;
    for (int l = 0; l < n; ++l)
      if (l != m) den *= X[m] - X[l];
    for (int q = 0; q < n; ++q) {
      if (q == m) continue;
      long long t = 1;
      for (int l = 0; l < n; ++l)
        if (l != m && l != q) t *= -X[l];
      num += t;
    }
    out[m] = ((double)(2 * num) / (double)den) / h;
  }
  return 0;
}

// H-AFDA one-sided weights (Kristek, Moczo & Archuleta 2002; the reference's draft
// Stencil.jl:52-105 formulas #1, #2, #4 and free-surface-kernels.jl:16-214), carried to every
// row an 8th-order centred stencil would take above the surface. Table rows: A0..A3 (d sxz/dz
// at z = k h from sxz(0) = 0 and sxz at (m+1/2) h), B0..B2 (d szz/dz at (k+1/2) h from szz at
// m h), C1..C3 (d vz/dz at k h from vz at (m+1/2) h), D0..D2 (d vx/dz at (k+1/2) h from vx at
// m h); column m multiplies grid row 4 + m, m < counts[row]. Each weight is the derivative of
// a Lagrange basis polynomial at x0,
//   w_m = sum_{q != m} prod_{l != m, q} (x0 - x_l) / prod_{l != m} (x_m - x_l),
// computed in integers of h/2 (every factor is at most 16, so both sums are exact in int64
// and in double): one rounding, in the final division.
static const int kAfdaDefault[13] = {3, 4, 6, 8, 3, 4, 6, 2, 4, 6, 3, 4, 6};

int vi_afda_default_points(int out[13]) {
  if (!out) return fail(VI_EINVAL, "vi_afda_default_points: null output");
  for (int r = 0; r < 13; ++r) out[r] = kAfdaDefault[r];
  return 0;
}

int vi_afda_weights(double h, const int* counts, double out[13][8]) {
  if (!out || !(h > 0)) return fail(VI_EINVAL, "vi_afda_weights: need h > 0 and an output");
  const int* nc = counts ? counts : kAfdaDefault;
  for (int r = 0; r < 13; ++r)
    if (nc[r] < (r < 4 ? 1 : 2) || nc[r] > 8)
      return fail(VI_EINVAL, "vi_afda_weights: point count " + std::to_string(r) +
                                 " outside [" + (r < 4 ? "1" : "2") + ", 8]");
  for (int r = 0; r < 13; ++r) {
    long long X[9];  // positions of the points in units of h/2
    long long X0;    // position of the derivative
    int np = 0, first = 0;
    const int n = nc[r];
    if (r < 4) {  // A_k: the boundary point first (value 0), its weight not stored
      X[np++] = 0;
      for (int m = 0; m < n; ++m) X[np++] = 2 * m + 1;
      first = 1;
      X0 = 2 * r;
    } else if (r < 7) {
      for (int m = 0; m < n; ++m) X[np++] = 2 * m;
      X0 = 2 * (r - 4) + 1;
    } else if (r < 10) {
      for (int m = 0; m < n; ++m) X[np++] = 2 * m + 1;
      X0 = 2 * (r - 6);
    } else {
      for (int m = 0; m < n; ++m) X[np++] = 2 * m;
      X0 = 2 * (r - 10) + 1;
    }
    for (int m = 0; m < 8; ++m) out[r][m] = 0.0;
    for (int m = first; m < np; ++m) {
      long long den = 1, num = 0;
      for (int l = 0; l < np; ++l)
        if (l != m) den *= X[m] - X[l];
      for (int q = 0; q < np; ++q) {
        if (q == m) continue;
        long long t = 1;
        for (int l = 0; l < np; ++l)
          if (l != m && l != q) t *= X0 - X[l];
        num += t;
      }
      // den holds one factor h/2 more than num: 2 num / den per unit h
      out[r][m - first] = ((double)(2 * num) / (double)den) / h;
    }
  }
  return 0;
}

// H-AFDA formula #3, the reference's Hermitian stencil (Stencil.jl:52-105 D3/D3_deriv from
// _stencil_hermitian, Stencil.jl:119-189): d vz/dz at z = h from d vz/dz(0) and vz at the half
// points (m+1/2) h, m < n, exact for polynomials of degree <= n. With the nodes x_m = 2m+1
// in units of h/2 and w(x) = prod (x - x_l), any such polynomial is sum p(x_m) l_m + c w, and
// p'(0) fixes c:
//   alpha = w'(2) / w'(0),   beta_m = l_m'(2) - alpha l_m'(0).
// In integers, W = w', N_m(x) = sum_{q != m} prod_{l != m, q} (x - x_l), D_m = prod_{l != m}
// (x_m - x_l): alpha = W(2)/W(0), beta_m = 2 (N_m(2) W(0) - W(2) N_m(0)) / (D_m W(0)) per unit
// h. For n <= 8 every integer is below 2^53, so each weight is rounded once, by its division.
int vi_afda_hermitian_weights(double h, int n, double out[9]) {
  if (!out || !(h > 0) || n < 2 || n > 8)
    return fail(VI_EINVAL, "vi_afda_hermitian_weights: need h > 0, 2 <= points <= 8, an output");
  long long X[8], N0[8], N2[8], W0 = 0, W2 = 0;
  for (int m = 0; m < n; ++m) X[m] = 2 * m + 1;
  for (int m = 0; m < n; ++m) {
    long long p0 = 1, p2 = 1;
    N0[m] = N2[m] = 0;
    for (int l = 0; l < n; ++l)
      if (l != m) {
        p0 *= -X[l];
        p2 *= 2 - X[l];
      }
    W0 += p0;
    W2 += p2;
    for (int q = 0; q < n; ++q) {
      if (q == m) continue;
      long long t0 = 1, t2 = 1;
      for (int l = 0; l < n; ++l)
        if (l != m && l != q) {
          t0 *= -X[l];
          t2 *= 2 - X[l];
        }
      N0[m] += t0;
      N2[m] += t2;
    }
  }
  for (int m = 0; m < 9; ++m) out[m] = 0.0;
  out[0] = (double)W2 / (double)W0;
  for (int m = 0; m < n; ++m) {
    long long D = 1;
    for (int l = 0; l < n; ++l)
      if (l != m) D *= X[m] - X[l];
    out[1 + m] = ((double)(2 * (N2[m] * W0 - W2 * N0[m])) / (double)(D * W0)) / h;
  }
  return 0;
}

double vi_cpml_auto_d0(int thickness, double vmax, double dx, double R, double power) {
  return -(power + 1.0) * vmax * std::log(R) / (2.0 * thickness * dx);  // CPML.jl:60
}

int vi_cpml_profile(int n, const vi_cpml_desc* c, double dt, int axis, int precision,
                    double* a, double* b, double* kappa) {
  if (!c || n <= 0 || !a || !b || !kappa || c->thickness < 0 || 2 * c->thickness > n ||
      (precision != VI_F32 && precision != VI_F64))
    return fail(VI_EINVAL, "vi_cpml_profile: bad arguments");
  const int lo = axis == 0 ? c->left : c->top;
  const int hi = axis == 0 ? c->right : c->bottom;
  if (precision == VI_F64) {
    cpml_axis(n, c->thickness, c->d0, c->alpha_max, c->kappa_max, c->power, dt, lo, hi, a, b,
              kappa);
    return 0;
  }
  std::vector<float> f(3 * (size_t)n);
  cpml_axis(n, c->thickness, c->d0, c->alpha_max, c->kappa_max, c->power, dt, lo, hi,
            f.data(), f.data() + n, f.data() + 2 * n);
  for (int i = 0; i < n; ++i) {
    a[i] = f[i];
    b[i] = f[n + i];
    kappa[i] = f[2 * n + i];
  }
  return 0;
}

int vi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int vi_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(VI_EINVAL, "vi_mem_info: null argument");
  VI_HIP(hipMemGetInfo(free_bytes, total_bytes));
  int dev = 0;
  VI_HIP(hipGetDevice(&dev));
  *free_bytes += spare::bytes_on(dev);  // kept for reuse, freed when an allocation needs it
  return 0;
}

int vi_cached_bytes(size_t* bytes) {
  if (!bytes) return fail(VI_EINVAL, "vi_cached_bytes: null argument");
  int dev = 0;
  VI_HIP(hipGetDevice(&dev));
  *bytes = spare::bytes_on(dev);
  return 0;
}

int vi_release_cached(void) {
  int dev = 0;
  VI_HIP(hipGetDevice(&dev));
  spare::drop(dev);
  return 0;
}

// ---------------------------------------------------------------- models
int vi_model_create(int nx, int nz, int precision, const void* rho, const void* lam,
                    const void* mu, const void* tau, vi_model* out) {
  if (!out || !rho || !lam || !mu) return fail(VI_EINVAL, "vi_model_create: null argument");
  *out = nullptr;
  if (nx < 16 || nz < 16) return fail(VI_EINVAL, "vi_model_create: nx and nz must be >= 16");
  if (precision != VI_F32 && precision != VI_F64)
    return fail(VI_EINVAL, "vi_model_create: bad precision");
  if (vi_device_count() < 1) return fail(VI_EDEVICE, "vi_model_create: no HIP device visible");
  hipStream_t s;
  VI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const void* src[6] = {rho, lam, mu, tau, nullptr, nullptr};
  std::shared_ptr<ModelStore> m;
  int rc = make_model_store(nx, nz, precision == VI_F32 ? 4 : 8, tau != nullptr, src, s, &m);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  *out = new vi_model_s{m};
  return 0;
}

int vi_model_create_vti(int nx, int nz, int precision, const void* rho, const void* c11,
                        const void* c13, const void* c33, const void* c55, const void* tau,
                        vi_model* out) {
  if (!out || !rho || !c11 || !c13 || !c33 || !c55)
    return fail(VI_EINVAL, "vi_model_create_vti: null argument");
  *out = nullptr;
  if (nx < 16 || nz < 16) return fail(VI_EINVAL, "vi_model_create_vti: nx and nz must be >= 16");
  if (precision != VI_F32 && precision != VI_F64)
    return fail(VI_EINVAL, "vi_model_create_vti: bad precision");
  if (vi_device_count() < 1) return fail(VI_EDEVICE, "vi_model_create_vti: no HIP device visible");
  hipStream_t s;
  VI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const void* src[6] = {rho, c13, c55, tau, c11, c33};
  std::shared_ptr<ModelStore> m;
  int rc = make_model_store(nx, nz, precision == VI_F32 ? 4 : 8, tau != nullptr, src, s, &m);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  *out = new vi_model_s{m};
  return 0;
}

int vi_model_destroy(vi_model m) {
  delete m;  // handles that use the model keep its arrays alive
  return 0;
}

// ---------------------------------------------------------------- single shot
int vi_create(const vi_grid_desc* g, vi_handle* out) {
  if (!g || !out) return fail(VI_EINVAL, "vi_create: null argument");
  *out = nullptr;
  if (g->nx < 16 || g->nz < 16) return fail(VI_EINVAL, "vi_create: nx and nz must be >= 16");
  if (g->precision != VI_F32 && g->precision != VI_F64)
    return fail(VI_EINVAL, "vi_create: precision must be VI_F32 or VI_F64");
  if (g->nrelax < 0 || g->nrelax > vi::MAXL)
    return fail(VI_EINVAL, "vi_create: nrelax must be in [0, 6]");
  for (int o : {g->fd_order_x, g->fd_order_z})
    if (o != 2 && o != 4 && o != 6 && o != 8)
      return fail(VI_EINVAL, "vi_create: fd_order_x and fd_order_z must be 2, 4, 6 or 8");
  if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0))
    return fail(VI_EINVAL, "vi_create: dx, dz, dt must be positive");
  if (vi_device_count() < 1) return fail(VI_EDEVICE, "vi_create: no HIP device visible");
  std::unique_ptr<vi_sim_s> h(new vi_sim_s());
  h->g = *g;
  h->L = g->nrelax;
  h->esz = g->precision == VI_F32 ? 4 : 8;
  VI_HIP(hipGetDevice(&h->device));
  VI_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  VI_HIP(hipEventCreateWithFlags(&h->own_ev, hipEventDisableTiming));
  VI_HIP(hipEventCreateWithFlags(&h->ext_ev, hipEventDisableTiming));
  h->ldx = (int)round_up((size_t)g->nx + 72, 64);
  h->nzp = g->nz + vi::NW * vi::RZ_MAX + 16;
  h->plane = (size_t)h->ldx * h->nzp;
  const size_t pb = h->plane * h->esz;
  if (pb >= (size_t(1) << 31))
    return fail(VI_EINVAL, "vi_create: one padded field plane must stay below 2 GiB");
  for (auto& b : h->state) VI_TRY(b.alloc(pb, h->stream));
  if (h->L) VI_TRY(h->M.alloc(3 * (size_t)h->L * pb, h->stream));
  VI_TRY(apply_tiling(h.get(), default_rz(g->nx, g->nz)));
  VI_HIP(hipStreamSynchronize(h->stream));
  *out = h.release();
  return 0;
}

int vi_destroy(vi_handle h) {
  if (!h) return 0;
  if (h->grad_busy) return fail(VI_ESTATE, "vi_destroy: end or cancel the gradient job first");
  (void)wait_external(h);
  (void)hipStreamSynchronize(h->stream);
  hipStream_t s = h->stream;
  hipEvent_t e0 = h->own_ev, e1 = h->ext_ev;
  delete h;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int vi_set_model(vi_handle h, const void* rho, const void* lam, const void* mu,
                 const void* tau) {
  VI_TRY(check_handle(h));
  if (!rho || !lam || !mu) return fail(VI_EINVAL, "vi_set_model: rho, lam, mu required");
  if (h->L && !tau) return fail(VI_EINVAL, "vi_set_model: tau required when nrelax > 0");
  VI_TRY(wait_external(h));
  const void* src[6] = {rho, lam, mu, h->L ? tau : nullptr, nullptr, nullptr};
  VI_HIP(hipStreamSynchronize(h->stream));
  std::shared_ptr<ModelStore> m;
  VI_TRY(make_model_store(h->g.nx, h->g.nz, h->esz, h->L > 0, src, h->stream, &m));
  h->model = m;
  h->model_x0 = 0;
  touch(h);
  return 0;
}

int vi_set_model_vti(vi_handle h, const void* rho, const void* c11, const void* c13,
                     const void* c33, const void* c55, const void* tau) {
  VI_TRY(check_handle(h));
  if (!rho || !c11 || !c13 || !c33 || !c55)
    return fail(VI_EINVAL, "vi_set_model_vti: rho, c11, c13, c33, c55 required");
  if (h->L && !tau) return fail(VI_EINVAL, "vi_set_model_vti: tau required when nrelax > 0");
  VI_TRY(wait_external(h));
  const void* src[6] = {rho, c13, c55, h->L ? tau : nullptr, c11, c33};
  VI_HIP(hipStreamSynchronize(h->stream));
  std::shared_ptr<ModelStore> m;
  VI_TRY(make_model_store(h->g.nx, h->g.nz, h->esz, h->L > 0, src, h->stream, &m));
  h->model = m;
  h->model_x0 = 0;
  touch(h);
  return 0;
}

int vi_set_model_window(vi_handle h, vi_model m, int x0) {
  VI_TRY(check_handle(h));
  if (!m || !m->m) return fail(VI_EINVAL, "vi_set_model_window: null model");
  const ModelStore& s = *m->m;
  if (s.esz != h->esz) return fail(VI_EINVAL, "vi_set_model_window: precision mismatch");
  if (s.nz != h->g.nz) return fail(VI_EINVAL, "vi_set_model_window: nz mismatch");
  if (x0 < 0 || x0 + h->g.nx > s.nx)
    return fail(VI_EINVAL, "vi_set_model_window: window [x0, x0+nx) outside the model");
  if (h->L && !s.has_tau) return fail(VI_EINVAL, "vi_set_model_window: model has no tau");
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->model = m->m;
  h->model_x0 = x0;
  touch(h);
  return 0;
}

int vi_set_relaxation(vi_handle h, const double* taun) {
  VI_TRY(check_handle(h));
  if (h->L == 0) return 0;
  if (!taun) return fail(VI_EINVAL, "vi_set_relaxation: null taun");
  for (int k = 0; k < h->L; ++k)
    if (!(taun[k] > 0)) return fail(VI_EINVAL, "vi_set_relaxation: taun must be positive");
  for (int k = 0; k < h->L; ++k) h->taun[k] = taun[k];
  h->relax_set = true;
  touch(h);
  return 0;
}

int vi_set_cpml(vi_handle h, const vi_cpml_desc* c) {
  VI_TRY(check_handle(h));
  VI_TRY(wait_external(h));
  if (!c) {
    h->cpml_on = false;
    touch(h);
    return 0;
  }
  if (c->thickness < 1 || 2 * c->thickness > h->g.nx || 2 * c->thickness > h->g.nz)
    return fail(VI_EINVAL, "vi_set_cpml: thickness must be in [1, min(nx, nz)/2]");
  if (!(c->d0 >= 0) || !(c->kappa_max >= 1) || !(c->alpha_max >= 0) || !(c->power >= 0))
    return fail(VI_EINVAL, "vi_set_cpml: need d0 >= 0, kappa_max >= 1, alpha_max >= 0");
  if (c->top && h->fs)
    return fail(VI_EINVAL, "vi_set_cpml: top strip conflicts with the free surface");
  const int nx = h->g.nx, nz = h->g.nz, T = c->thickness;
  // [ax bx kx (nx each) | az bz kz (nz each)] in the handle's precision
  std::vector<char> hostc((size_t)(3 * nx + 3 * nz) * h->esz);
  auto fill = [&](auto* f) {
    cpml_axis(nx, T, c->d0, c->alpha_max, c->kappa_max, c->power, h->g.dt, c->left, c->right,
              f, f + nx, f + 2 * nx);
    cpml_axis(nz, T, c->d0, c->alpha_max, c->kappa_max, c->power, h->g.dt, c->top, c->bottom,
              f + 3 * nx, f + 3 * nx + nz, f + 3 * nx + 2 * nz);
  };
  if (h->esz == 4) fill((float*)hostc.data());
  else fill((double*)hostc.data());
  VI_HIP(hipStreamSynchronize(h->stream));
  VI_TRY(h->coef.alloc(hostc.size(), h->stream));
  VI_HIP(hipMemcpyAsync(h->coef.p, hostc.data(), hostc.size(), hipMemcpyHostToDevice,
                        h->stream));
  for (int q = 0; q < 4; ++q) VI_TRY(h->psi[q].alloc((size_t)nz * 2 * T * h->esz, h->stream));
  for (int q = 4; q < 8; ++q) VI_TRY(h->psi[q].alloc((size_t)2 * T * h->ldx * h->esz, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->cpml = *c;
  h->cpml_on = true;
  touch(h);
  return 0;
}

// The H-AFDA table of a handle: the weights [13][8], then for the adjoint gathers the
// difference G_a - G_c between each one-sided row j and the centred stencil it replaces,
// over the target rows g = 4..11 (G_c of D- at j: D[g-j] for 0 <= g-j <= 3, -D[j-g-1] for
// 0 <= j-g-1 <= 3; of D+: D[g-j-1], -D[j-g]), then formula #3: beta_m / h (8), alpha.
// mode: the free-surface mode the table is for (2 or 3); with 3 the adjoint's difference row
// of source row 5 is the Hermitian formula's vz weights minus the centred D-
static int upload_afda(vi_sim_s* h, int mode) {
  double w[13][8], D[4], hw[9];
  VI_TRY(vi_afda_weights(h->g.dz, h->afda_pts, w));
  VI_TRY(vi_afda_hermitian_weights(h->g.dz, h->herm_pts, hw));
  vi_fd_coeffs(h->g.dz, D);
  auto gc = [&](bool minus, int j, int g) -> double {
    const int a = minus ? g - j : g - j - 1, b = minus ? j - g - 1 : j - g;
    if (a >= 0 && a <= 3) return D[a];
    if (b >= 0 && b <= 3) return -D[b];
    return 0.0;
  };
  double c[13][8];
  for (int r = 0; r < 13; ++r) {
    // source row and operator: A_k D- at 4+k, B_k D+ at 4+k, C_k D- at 4+k (k = 1..3),
    // D_k D+ at 4+k
    const int j = r < 4 ? 4 + r : r < 7 ? r : r < 10 ? r - 2 : r - 6;
    const bool minus = r < 4 || (r >= 7 && r < 10);
    for (int m = 0; m < 8; ++m) c[r][m] = w[r][m] - gc(minus, j, 4 + m);
  }
  if (mode == 3)  // row 5's d vz/dz = alpha es + sum_m beta_m vz(4 + m) (C1's place)
    for (int m = 0; m < 8; ++m) c[7][m] = hw[1 + m] - gc(true, 5, 4 + m);
  std::vector<char> t(224 * h->esz);  // 208 + 8 + 1, rounded up
  auto put = [&](int at, double v) {
    if (h->esz == 4) ((float*)t.data())[at] = (float)v;
    else ((double*)t.data())[at] = v;
  };
  for (int r = 0; r < 13; ++r)
    for (int m = 0; m < 8; ++m) {
      put(r * 8 + m, w[r][m]);
      put(104 + r * 8 + m, c[r][m]);
    }
  for (int m = 0; m < 8; ++m) put(208 + m, hw[1 + m]);
  put(216, hw[0]);
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  VI_TRY(h->afda.alloc(t.size(), h->stream));
  VI_HIP(hipMemcpyAsync(h->afda.p, t.data(), t.size(), hipMemcpyHostToDevice, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int vi_set_free_surface(vi_handle h, int mode) {
  VI_TRY(check_handle(h));
  if (mode < 0 || mode > 3)
    return fail(VI_EINVAL, "vi_set_free_surface: mode must be 0 (none), 1 (imaging), 2 (H-AFDA) "
                           "or 3 (H-AFDA with formula #3)");
  if (mode && h->cpml_on && h->cpml.top)
    return fail(VI_EINVAL, "vi_set_free_surface: conflicts with a top CPML strip");
  if (mode && h->g.nz < 17) return fail(VI_EINVAL, "vi_set_free_surface: nz must be >= 17");
  if (mode >= 2 && h->g.fd_order_z != 8)
    return fail(VI_EINVAL, "vi_set_free_surface: the H-AFDA rows need fd_order_z = 8");
  if (mode >= 2) VI_TRY(upload_afda(h, mode));
  h->fs = mode;
  touch(h);
  return 0;
}

int vi_set_afda_points(vi_handle h, const int* counts) {
  VI_TRY(check_handle(h));
  int pts[13];
  for (int r = 0; r < 13; ++r) pts[r] = counts ? counts[r] : kAfdaDefault[r];
  double w[13][8];
  VI_TRY(vi_afda_weights(1.0, pts, w));  // range check
  std::memcpy(h->afda_pts, pts, sizeof pts);
  if (h->fs >= 2) VI_TRY(upload_afda(h, h->fs));
  touch(h);
  return 0;
}

int vi_set_afda_hermitian(vi_handle h, int points) {
  VI_TRY(check_handle(h));
  double w[9];
  VI_TRY(vi_afda_hermitian_weights(1.0, points, w));  // range check
  h->herm_pts = points;
  if (h->fs >= 2) VI_TRY(upload_afda(h, h->fs));
  touch(h);
  return 0;
}

int vi_set_sources(vi_handle h, int n, const int* ix, const int* iz, const int* kind, int nt,
                   const void* wavelets) {
  VI_TRY(check_handle(h));
  if (n < 0 || nt < 0 || (n > 0 && (!ix || !iz || !kind || (nt > 0 && !wavelets))))
    return fail(VI_EINVAL, "vi_set_sources: bad arguments");
  for (int q = 0; q < n; ++q) {
    if (!interior(h, ix[q], iz[q]))
      return fail(VI_EINVAL, "vi_set_sources: source " + std::to_string(q) +
                                 " is outside the interior [4, nx-4) x [4, nz-4)");
    if (kind[q] < 0 || kind[q] > 2) return fail(VI_EINVAL, "vi_set_sources: bad kind");
  }
  const size_t wbytes = (size_t)n * nt * h->esz;
  if (wbytes >= (size_t(1) << 31))
    return fail(VI_EINVAL, "vi_set_sources: the wavelet table must stay below 2 GiB");
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->src_ix.assign(ix, ix + n);
  h->src_iz.assign(iz, iz + n);
  h->src_kind.assign(kind, kind + n);
  h->nsrc = n;
  h->snt = nt;
  h->src_w_host.assign((const char*)wavelets, (const char*)wavelets + (n && nt ? wbytes : 0));
  std::vector<int> order;
  VI_TRY(upload_table(h, h->src_tab, n, ix, iz, kind, &order));
  VI_TRY(upload_src_wc(h, order));
  VI_TRY(h->src_w.alloc(wbytes, h->stream));
  if (n && nt)
    VI_HIP(hipMemcpyAsync(h->src_w.p, wavelets, wbytes, hipMemcpyHostToDevice, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  touch(h);
  return 0;
}

int vi_set_receivers(vi_handle h, int n, const int* ix, const int* iz, const int* comp,
                     int nt_max) {
  VI_TRY(check_handle(h));
  if (n < 0 || nt_max < 0 || (n > 0 && (!ix || !iz || !comp)))
    return fail(VI_EINVAL, "vi_set_receivers: bad arguments");
  for (int q = 0; q < n; ++q) {
    if (!interior(h, ix[q], iz[q]))
      return fail(VI_EINVAL, "vi_set_receivers: receiver " + std::to_string(q) +
                                 " is outside the interior [4, nx-4) x [4, nz-4)");
    if (comp[q] < 0 || comp[q] > 4) return fail(VI_EINVAL, "vi_set_receivers: bad component");
  }
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->rec_ix.assign(ix, ix + n);
  h->rec_iz.assign(iz, iz + n);
  h->rec_comp.assign(comp, comp + n);
  h->nrec = n;
  VI_TRY(upload_table(h, h->rec_tab, n, ix, iz, comp));
  VI_TRY(h->traces.alloc((size_t)n * nt_max * h->esz, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->rec_nt = nt_max;
  touch(h);
  return 0;
}

int vi_set_state(vi_handle h, int field, const void* host) {
  VI_TRY(check_handle(h));
  if (!host) return fail(VI_EINVAL, "vi_set_state: null buffer");
  const size_t n = field_count(h, field);
  if (!n) return fail(VI_EINVAL, "vi_set_state: field not present in this configuration");
  VI_TRY(wait_external(h));
  if (field <= VI_SXZ) {
    VI_TRY(upload_plane(h->stream, h->state[field].p, h->ldx, host, h->g.nx, h->g.nz, h->esz));
  } else if (field <= VI_MXZ) {
    const size_t nc = (size_t)h->g.nx * h->g.nz;
    for (int k = 0; k < h->L; ++k) {
      char* dst = (char*)h->M.p + ((size_t)(field - VI_MXX) * h->L + k) * h->plane * h->esz;
      VI_TRY(upload_plane(h->stream, dst, h->ldx, (const char*)host + k * nc * h->esz, h->g.nx,
                          h->g.nz, h->esz));
    }
  } else {
    const int q = field - VI_PSI_VX_X;
    if (q < 4) {
      VI_HIP(hipMemcpyAsync(h->psi[q].p, host, n * h->esz, hipMemcpyHostToDevice, h->stream));
    } else {
      VI_TRY(upload_plane(h->stream, h->psi[q].p, h->ldx, host, h->g.nx,
                          2 * h->cpml.thickness, h->esz));
    }
  }
  VI_HIP(hipStreamSynchronize(h->stream));
  h->own_pending = false;
  return 0;
}

int vi_get_state(vi_handle h, int field, void* host) {
  VI_TRY(check_handle(h));
  if (!host) return fail(VI_EINVAL, "vi_get_state: null buffer");
  const size_t n = field_count(h, field);
  if (!n) return fail(VI_EINVAL, "vi_get_state: field not present in this configuration");
  VI_TRY(wait_external(h));
  if (field <= VI_SXZ) {
    VI_TRY(download_plane(h->stream, host, h->state[field].p, h->ldx, h->g.nx, h->g.nz, h->esz));
  } else if (field <= VI_MXZ) {
    const size_t nc = (size_t)h->g.nx * h->g.nz;
    for (int k = 0; k < h->L; ++k) {
      const char* src =
          (const char*)h->M.p + ((size_t)(field - VI_MXX) * h->L + k) * h->plane * h->esz;
      VI_TRY(download_plane(h->stream, (char*)host + k * nc * h->esz, src, h->ldx, h->g.nx,
                            h->g.nz, h->esz));
    }
  } else {
    const int q = field - VI_PSI_VX_X;
    if (q < 4) {
      VI_HIP(hipMemcpyAsync(host, h->psi[q].p, n * h->esz, hipMemcpyDeviceToHost, h->stream));
    } else {
      VI_TRY(download_plane(h->stream, host, h->psi[q].p, h->ldx, h->g.nx,
                            2 * h->cpml.thickness, h->esz));
    }
  }
  VI_HIP(hipStreamSynchronize(h->stream));
  h->own_pending = false;
  return 0;
}

int vi_reset(vi_handle h) {
  VI_TRY(check_handle(h));
  VI_TRY(wait_external(h));
  for (auto* b : mutable_arrays(h))
    if (b->p) VI_HIP(hipMemsetAsync(b->p, 0, b->bytes, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->own_pending = false;
  h->step = 0;
  touch(h);
  return 0;
}

int vi_step(vi_handle h, int nsteps) {
  VI_TRY(check_handle(h));
  if (nsteps < 0) return fail(VI_EINVAL, "vi_step: nsteps < 0");
  if (nsteps == 0) return check_ready(h);
  return step_single(h, nsteps);
}

int vi_sync(vi_handle h) {
  VI_TRY(check_handle(h));
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->own_pending = false;
  return 0;
}

int vi_rates(vi_handle h, const void* src_values, void* dv, void* ds, void* dm) {
  VI_TRY(rates_checks(h, dv, ds, "vi_rates"));
  VI_TRY(h->rates_ws.reserve(rates_bytes(h)));
  return h->esz == 4 ? rates_run<float>(h, src_values, dv, ds, dm, (char*)h->rates_ws.p)
                     : rates_run<double>(h, src_values, dv, ds, dm, (char*)h->rates_ws.p);
}

int vi_rates_noalloc(vi_handle h, const void* src_values, void* dv, void* ds, void* dm) {
  VI_TRY(rates_checks(h, dv, ds, "vi_rates_noalloc"));
  const size_t need = rates_bytes(h);
  if (need > kRatesScratchBytes)
    return fail(VI_EINVAL, "vi_rates_noalloc: this handle needs " + std::to_string(need) +
                               " bytes of workspace, more than the " +
                               std::to_string(kRatesScratchBytes) + " of the static one");
  char* ws = nullptr;
  VI_TRY(scratch_on(h->device, &ws));
  return h->esz == 4 ? rates_run<float>(h, src_values, dv, ds, dm, ws)
                     : rates_run<double>(h, src_values, dv, ds, dm, ws);
}

int vi_trim(vi_handle h) {
  VI_TRY(check_handle(h));
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->grad_ck.release();
  h->grad_seg.release();
  h->grad_adj.release();
  h->grad_host.release();
  h->rates_ws.release();
  return 0;
}

int vi_check_finite(vi_handle h, int* finite) {
  VI_TRY(check_handle(h));
  if (!finite) return fail(VI_EINVAL, "vi_check_finite: null argument");
  VI_TRY(wait_external(h));
  hipStream_t st = h->stream;
  VI_TRY(h->check_flag.reserve(256));
  int flag = 0;
  VI_HIP(hipMemsetAsync(h->check_flag.p, 0, sizeof(int), st));
  const dim3 grid((h->g.nx - 8 + 63) / 64, (h->g.nz - 8 + 3) / 4), block(64, 4);
  if (h->esz == 4)
    hipLaunchKernelGGL(vi::k_nonfinite<float>, grid, block, 0, st, (const float*)h->state[0].p,
                       (const float*)h->state[1].p, h->g.nx, h->g.nz, h->ldx,
                       (int*)h->check_flag.p);
  else
    hipLaunchKernelGGL(vi::k_nonfinite<double>, grid, block, 0, st, (const double*)h->state[0].p,
                       (const double*)h->state[1].p, h->g.nx, h->g.nz, h->ldx,
                       (int*)h->check_flag.p);
  VI_HIP(hipGetLastError());
  VI_HIP(hipMemcpyAsync(&flag, h->check_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  VI_HIP(hipStreamSynchronize(st));
  h->own_pending = false;
  *finite = flag ? 0 : 1;
  return 0;
}

long vi_step_index(vi_handle h) { return h ? h->step : -1; }

int vi_get_traces(vi_handle h, void* out, int nt) {
  VI_TRY(check_handle(h));
  if (nt < 0 || nt > h->rec_nt || (!out && nt && h->nrec))
    return fail(VI_EINVAL, "vi_get_traces: nt exceeds the recorded capacity");
  if (!h->nrec || !nt) return 0;
  VI_TRY(wait_external(h));
  std::vector<char> tmp((size_t)nt * h->nrec * h->esz);
  VI_HIP(hipMemcpyAsync(tmp.data(), h->traces.p, tmp.size(), hipMemcpyDeviceToHost, h->stream));
  VI_HIP(hipStreamSynchronize(h->stream));
  h->own_pending = false;
  for (int r = 0; r < h->nrec; ++r)  // device [t][rec] -> host [rec][t]
    for (int t = 0; t < nt; ++t)
      std::memcpy((char*)out + ((size_t)r * nt + t) * h->esz,
                  tmp.data() + ((size_t)t * h->nrec + r) * h->esz, h->esz);
  return 0;
}

int vi_device_bytes(vi_handle h, size_t* bytes) {
  VI_TRY(check_handle(h));
  if (!bytes) return fail(VI_EINVAL, "vi_device_bytes: null argument");
  size_t t = 0;
  for (auto* b : mutable_arrays(h)) t += b->bytes;
  t += h->coef.bytes + h->src_tab.bytes + h->src_w.bytes + h->src_wc.bytes + h->rec_tab.bytes;
  t += h->desc_dev.bytes + h->tile_shot_dev.bytes + h->rates_ws.bytes;
  t += h->grad_ck.bytes + h->grad_seg.bytes + h->grad_adj.bytes + h->check_flag.bytes;
  if (h->model && h->model.use_count() == 1)
    for (auto& b : h->model->arr) t += b.bytes;
  *bytes = t;
  return 0;
}

int vi_gradient_bytes(vi_handle h, int nt, int checkpoint_every, size_t* bytes) {
  VI_TRY(check_handle(h));
  if (!bytes || nt < 1 || checkpoint_every < 0)
    return fail(VI_EINVAL, "vi_gradient_bytes: need nt >= 1, checkpoint_every >= 0, bytes");
  std::vector<DevBuf*> live = mutable_arrays(h);
  size_t lb[14];
  for (int q = 0; q < 14; ++q) lb[q] = live[q]->bytes;
  const size_t ints = 4 * (size_t)h->nrec + 1;  // bound: targets <= nrec
  *bytes = grad_layout(nt, checkpoint_every, h->L, h->plane * h->esz, h->ldx, h->esz, lb,
                       h->nrec, ints).total();
  return 0;
}

int vi_plan_bytes(const vi_grid_desc* g, int vti, int cpml_thickness, int nsrc, int src_nt,
                  int nrec, int rec_nt, int grad_nt, int checkpoint_every, size_t* bytes) {
  if (!g || !bytes) return fail(VI_EINVAL, "vi_plan_bytes: null argument");
  if (g->nx < 16 || g->nz < 16 || (g->precision != VI_F32 && g->precision != VI_F64) ||
      g->nrelax < 0 || g->nrelax > vi::MAXL)
    return fail(VI_EINVAL, "vi_plan_bytes: bad grid description");
  const int T = cpml_thickness;
  if (T < 0 || 2 * T > g->nx || 2 * T > g->nz || nsrc < 0 || src_nt < 0 || nrec < 0 ||
      rec_nt < 0 || grad_nt < 0 || checkpoint_every < 0)
    return fail(VI_EINVAL, "vi_plan_bytes: negative count or CPML wider than the grid");
  const size_t esz = g->precision == VI_F32 ? 4 : 8;
  const size_t nx = g->nx, nz = g->nz, L = g->nrelax;
  const size_t ldx = round_up(nx + 72, 64), nzp = nz + vi::NW * vi::RZ_MAX + 16;
  const size_t pb = ldx * nzp * esz;
  size_t live[14] = {pb, pb, pb, pb, pb, 3 * L * pb};
  for (int q = 0; q < 8; ++q)
    live[6 + q] = T ? (q < 4 ? nz * 2 * T * esz : 2 * (size_t)T * ldx * esz) : 0;
  size_t t = 0;
  for (size_t b : live) t += b;
  t += (3 + (L ? 1 : 0) + (vti ? 2 : 0)) * pb;  // model planes (same padding as the state)
  if (T) t += (3 * nx + 3 * nz) * esz;           // CPML a, b, kappa
  // tile tables at the finest tiling, descriptors, source wavelets and traces
  const size_t ntiles = ((nx - 8 + vi::TX - 1) / vi::TX) * ((nz - 8 + vi::NW - 1) / vi::NW);
  t += (2 * (ntiles + 1) + 3 * ((size_t)nsrc + nrec)) * sizeof(int) + ntiles * sizeof(int);
  t += std::max(desc_size(4), desc_size(8));
  t += 2 * (size_t)nsrc * src_nt * esz + (size_t)nrec * rec_nt * esz;  // wavelets twice
  if (grad_nt > 0)
    t += grad_layout(grad_nt, checkpoint_every, (int)L, pb, ldx, esz, live, nrec,
                     4 * (size_t)nrec + 1).total();
  *bytes = t;
  return 0;
}

int vi_set_tiling(vi_handle h, int rz) {
  VI_TRY(check_handle(h));
  if (rz != 1 && rz != 2 && rz != 4 && rz != 8)
    return fail(VI_EINVAL, "vi_set_tiling: rows per wavefront must be 1, 2, 4 or 8");
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  return apply_tiling(h, rz);
}

int vi_get_tiling(vi_handle h) { return h ? h->rz : -1; }

int vi_autotune(vi_handle h, int trial_steps, int* chosen, double* us_per_step) {
  VI_TRY(check_handle(h));
  if (trial_steps < 1) return fail(VI_EINVAL, "vi_autotune: trial_steps must be >= 1");
  VI_TRY(check_ready(h));
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  // snapshot every mutable array so that the trials leave the state untouched
  std::vector<DevBuf*> live = mutable_arrays(h);
  std::vector<DevBuf> snap(live.size());
  for (size_t q = 0; q < live.size(); ++q) {
    if (!live[q]->p) continue;
    VI_TRY(snap[q].alloc(live[q]->bytes, h->stream));
    VI_HIP(hipMemcpyAsync(snap[q].p, live[q]->p, live[q]->bytes, hipMemcpyDeviceToDevice,
                          h->stream));
  }
  const long step0 = h->step;
  const bool prof0 = h->prof.on;
  h->prof.on = false;
  hipEvent_t e0, e1;
  VI_HIP(hipEventCreate(&e0));
  VI_HIP(hipEventCreate(&e1));
  const int cands[4] = {1, 2, 4, 8};
  int best = h->rz;
  double best_us = 1e300;
  int rc = 0;
  for (int c = 0; c < 4 && !rc; ++c) {
    if ((rc = apply_tiling(h, cands[c]))) break;
    if ((rc = step_single(h, 2))) break;
    if ((rc = hipEventRecord(e0, h->stream) == hipSuccess ? 0 : VI_EDEVICE)) break;
    if ((rc = step_single(h, trial_steps))) break;
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
    for (size_t q = 0; q < live.size(); ++q)
      if (live[q]->p && hipMemcpyAsync(live[q]->p, snap[q].p, live[q]->bytes,
                                       hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
        rc = VI_EDEVICE;
    h->step = step0;
    touch(h);
  }
  (void)hipStreamSynchronize(h->stream);
  h->step = step0;
  h->prof.on = prof0;
  h->own_pending = false;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return g_err.empty() ? fail(rc, "vi_autotune: HIP error") : rc;
  VI_TRY(apply_tiling(h, best));
  if (chosen) *chosen = best;
  return 0;
}

int vi_profile_enable(vi_handle h, int every) {
  VI_TRY(check_handle(h));
  if (every < 0) return fail(VI_EINVAL, "vi_profile_enable: every must be >= 0");
  VI_TRY(h->prof.collect(h->stream));
  h->prof.on = every != 0;
  h->prof.every = every ? every : 1;
  h->prof.reset();
  return 0;
}

int vi_profile_read(vi_handle h, int kernel, double* ms_total, long* launches) {
  VI_TRY(check_handle(h));
  if (kernel < 0 || kernel > 1 || !ms_total || !launches)
    return fail(VI_EINVAL, "vi_profile_read: bad arguments");
  VI_TRY(h->prof.collect(h->stream));
  *ms_total = h->prof.ms[kernel];
  *launches = h->prof.n[kernel];
  return 0;
}

int vi_algorithmic_bytes(vi_handle h, int kernel, double* bytes) {
  VI_TRY(check_handle(h));
  if (kernel < 0 || kernel > 1 || !bytes) return fail(VI_EINVAL, "vi_algorithmic_bytes");
  const double ni = (double)(h->g.nx - 8) * (h->g.nz - 8);
  // velocity: v R+W (4) + s R (3) + rho (1); stress: v R (2) + s R+W (6) + lam, mu (2)
  // [+ tau (1) + M R+W (6L)]
  double words = kernel == 0 ? 8.0 : (h->L ? 11.0 + 6.0 * h->L : 10.0);
  if (kernel == 1 && h->model && h->model->vti) words += 2.0;  // C11, C33
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


static int gradient_checks(vi_handle h, int nt, const void* dobs, int checkpoint_every,
                           const char* who) {
  VI_TRY(check_handle(h));
  if (nt < 1 || !dobs || checkpoint_every < 0)
    return fail(VI_EINVAL, std::string(who) + ": need nt >= 1, dobs, checkpoint_every >= 0");
  VI_TRY(check_ready(h));
  if (h->nrec < 1 || h->rec_nt < nt)
    return fail(VI_ESTATE, std::string(who) + ": receivers with capacity >= nt are required");
  VI_TRY(wait_external(h));
  VI_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

static int gradient_run(vi_handle h, int nt, const void* dobs, int checkpoint_every,
                        double* misfit, void* const out[6]) {
  vi_grad g = nullptr;
  VI_TRY(gradient_begin(h, nt, dobs, checkpoint_every, &g));
  std::unique_ptr<vi_grad_s> own(g);
  const int rc = g->finish(misfit, out);
  if (rc) h->grad_busy = false;
  return rc;
}

int vi_gradient(vi_handle h, int nt, const void* dobs, int checkpoint_every, double* misfit,
                void* g_rho, void* g_lam, void* g_mu, void* g_tau) {
  VI_TRY(gradient_checks(h, nt, dobs, checkpoint_every, "vi_gradient"));
  if (h->model->vti) return fail(VI_ESTATE, "vi_gradient: VTI model, use vi_gradient_vti");
  void* out[6] = {g_rho, g_lam, g_mu, g_tau, nullptr, nullptr};
  return gradient_run(h, nt, dobs, checkpoint_every, misfit, out);
}

int vi_gradient_vti(vi_handle h, int nt, const void* dobs, int checkpoint_every,
                    double* misfit, void* g_rho, void* g_c11, void* g_c13, void* g_c33,
                    void* g_c55, void* g_tau) {
  VI_TRY(gradient_checks(h, nt, dobs, checkpoint_every, "vi_gradient_vti"));
  if (!h->model->vti) return fail(VI_ESTATE, "vi_gradient_vti: isotropic model, use vi_gradient");
  void* out[6] = {g_rho, g_c13, g_c55, g_tau, g_c11, g_c33};
  return gradient_run(h, nt, dobs, checkpoint_every, misfit, out);
}

int vi_gradient_begin(vi_handle h, int nt, const void* dobs, int checkpoint_every,
                      vi_grad* out) {
  if (!out) return fail(VI_EINVAL, "vi_gradient_begin: null output");
  *out = nullptr;
  VI_TRY(gradient_checks(h, nt, dobs, checkpoint_every, "vi_gradient_begin"));
  return gradient_begin(h, nt, dobs, checkpoint_every, out);
}

int vi_gradient_advance(vi_grad g, int max_steps, int* finished) {
  if (!g || max_steps < 1) return fail(VI_EINVAL, "vi_gradient_advance: bad arguments");
  VI_TRY(g->advance(max_steps));
  if (finished) *finished = g->done ? 1 : 0;
  return 0;
}

int vi_gradient_idle(vi_grad g) {
  if (!g) return fail(VI_EINVAL, "vi_gradient_idle: null job");
  const hipError_t e = hipEventQuery(g->ev);
  if (e == hipErrorNotReady) return 0;
  VI_HIP(e);
  return 1;
}

int vi_gradient_end(vi_grad g, double* misfit, void* g_rho, void* g_lam_or_c13,
                    void* g_mu_or_c55, void* g_tau, void* g_c11, void* g_c33) {
  if (!g) return fail(VI_EINVAL, "vi_gradient_end: null job");
  std::unique_ptr<vi_grad_s> own(g);
  void* out[6] = {g_rho, g_lam_or_c13, g_mu_or_c55, g_tau, g_c11, g_c33};
  const int rc = g->finish(misfit, out);
  if (rc) g->h->grad_busy = false;
  return rc;
}

int vi_gradient_cancel(vi_grad g) {
  if (!g) return 0;
  std::unique_ptr<vi_grad_s> own(g);
  vi_sim_s* h = g->h;
  (void)hipStreamSynchronize(h->stream);
  h->grad_busy = false;
  return vi_reset(h);
}

// ---------------------------------------------------------------- batches
int vi_batch_create(vi_batch* out) {
  if (!out) return fail(VI_EINVAL, "vi_batch_create: null argument");
  *out = nullptr;
  if (vi_device_count() < 1) return fail(VI_EDEVICE, "vi_batch_create: no HIP device visible");
  std::unique_ptr<vi_batch_s> b(new vi_batch_s());
  VI_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  VI_HIP(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
  for (int k = 0; k < vi_batch_s::MAX_STREAMS - 1; ++k) {
    VI_HIP(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking));
    VI_HIP(hipEventCreateWithFlags(&b->join_ev[k], hipEventDisableTiming));
  }
  *out = b.release();
  return 0;
}

int vi_batch_destroy(vi_batch b) {
  if (!b) return 0;
  (void)hipStreamSynchronize(b->stream);
  for (auto s : b->side)
    if (s) (void)hipStreamSynchronize(s);
  invalidate_graph(b);
  constexpr int NS = vi_batch_s::MAX_STREAMS - 1;
  hipStream_t s = b->stream, side[NS];
  hipEvent_t e0 = b->fork_ev, ej[NS];
  for (int k = 0; k < NS; ++k) side[k] = b->side[k], ej[k] = b->join_ev[k];
  delete b;
  if (e0) (void)hipEventDestroy(e0);
  for (int k = 0; k < NS; ++k) {
    if (ej[k]) (void)hipEventDestroy(ej[k]);
    if (side[k]) (void)hipStreamDestroy(side[k]);
  }
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int vi_batch_set_streams(vi_batch b, int n) {
  if (!b || n < 1 || n > vi_batch_s::MAX_STREAMS)
    return fail(VI_EINVAL, "vi_batch_set_streams: n must be 1 to 4");
  VI_HIP(hipStreamSynchronize(b->stream));
  b->nstreams = n;
  b->built = false;
  invalidate_graph(b);
  return 0;
}

int vi_batch_set(vi_batch b, const vi_handle* hs, int n) {
  if (!b || n < 0 || (n && !hs)) return fail(VI_EINVAL, "vi_batch_set: bad arguments");
  for (int q = 0; q < n; ++q) {
    if (!hs[q]) return fail(VI_EINVAL, "vi_batch_set: null handle");
    if (hs[q]->grad_busy)
      return fail(VI_ESTATE, "vi_batch_set: handle in use by a gradient job");
    for (int p = 0; p < q; ++p)
      if (hs[p] == hs[q]) return fail(VI_EINVAL, "vi_batch_set: handle listed twice");
  }
  VI_HIP(hipStreamSynchronize(b->stream));
  b->members.assign(hs, hs + n);
  b->built = false;
  invalidate_graph(b);
  return 0;
}

int vi_batch_set_tiling(vi_batch b, int rz) {
  if (!b) return fail(VI_EINVAL, "null batch");
  if (rz != 1 && rz != 2 && rz != 4 && rz != 8)
    return fail(VI_EINVAL, "vi_batch_set_tiling: rows per wavefront must be 1, 2, 4 or 8");
  VI_HIP(hipStreamSynchronize(b->stream));
  b->rz = rz;
  b->built = false;
  invalidate_graph(b);
  return 0;
}

int vi_batch_get_tiling(vi_batch b) { return b ? b->rz : -1; }

size_t vi_batch_get_cache_block(vi_batch b) { return b ? b->block_bytes : 0; }

int vi_batch_set_cache_block(vi_batch b, size_t bytes) {
  if (!b) return fail(VI_EINVAL, "null batch");
  VI_HIP(hipStreamSynchronize(b->stream));
  b->block_bytes = bytes;
  b->built = false;
  invalidate_graph(b);
  return 0;
}

int vi_batch_set_graph(vi_batch b, int steps_per_graph) {
  if (!b || steps_per_graph < 0) return fail(VI_EINVAL, "vi_batch_set_graph: bad arguments");
  VI_HIP(hipStreamSynchronize(b->stream));
  invalidate_graph(b);
  b->graph_steps = steps_per_graph;
  return 0;
}

int vi_batch_step(vi_batch b, int nsteps) {
  if (!b || nsteps < 0) return fail(VI_EINVAL, "vi_batch_step: bad arguments");
  if (b->members.empty() || nsteps == 0) return 0;
  VI_TRY(batch_prepare(b));
  int left = nsteps;
  if (b->graph_steps > 0 && !b->prof.on) {
    const int K = b->graph_steps;
    if (left >= K && !b->graph) {
      VI_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = batch_launch(b, K);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(VI_EDEVICE, "vi_batch_step: graph capture failed");
      b->graph_def = g;
      VI_HIP(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
    }
    while (left >= K) {
      VI_HIP(hipGraphLaunch(b->graph, b->stream));
      left -= K;
    }
  }
  if (left) VI_TRY(batch_launch(b, left));
  return batch_finish(b, nsteps);
}

int vi_batch_sync(vi_batch b) {
  if (!b) return fail(VI_EINVAL, "null batch");
  VI_HIP(hipStreamSynchronize(b->stream));
  return 0;
}

int vi_batch_autotune(vi_batch b, int trial_steps, int* chosen, double* us_per_step) {
  if (!b || trial_steps < 1) return fail(VI_EINVAL, "vi_batch_autotune: bad arguments");
  if (b->members.empty()) return fail(VI_ESTATE, "vi_batch_autotune: empty batch");
  VI_TRY(batch_prepare(b));
  VI_HIP(hipStreamSynchronize(b->stream));
  for (auto* h : b->members) VI_TRY(wait_external(h));
  std::vector<std::vector<DevBuf>> snaps(b->members.size());
  std::vector<long> step0(b->members.size());
  for (size_t m = 0; m < b->members.size(); ++m) {
    vi_sim_s* h = b->members[m];
    step0[m] = h->step;
    auto live = mutable_arrays(h);
    snaps[m] = std::vector<DevBuf>(live.size());
    for (size_t q = 0; q < live.size(); ++q) {
      if (!live[q]->p) continue;
      VI_TRY(snaps[m][q].alloc(live[q]->bytes, b->stream));
      VI_HIP(hipMemcpyAsync(snaps[m][q].p, live[q]->p, live[q]->bytes,
                            hipMemcpyDeviceToDevice, b->stream));
    }
  }
  auto restore = [&]() -> int {
    for (size_t m = 0; m < b->members.size(); ++m) {
      vi_sim_s* h = b->members[m];
      auto live = mutable_arrays(h);
      for (size_t q = 0; q < live.size(); ++q)
        if (live[q]->p)
          VI_HIP(hipMemcpyAsync(live[q]->p, snaps[m][q].p, live[q]->bytes,
                                hipMemcpyDeviceToDevice, b->stream));
      h->step = step0[m];
      touch(h);
    }
    return 0;
  };
  const bool prof0 = b->prof.on;
  b->prof.on = false;
  const int graph0 = b->graph_steps;
  b->graph_steps = 0;
  hipEvent_t e0, e1;
  VI_HIP(hipEventCreate(&e0));
  VI_HIP(hipEventCreate(&e1));
  // one trial of the current tiling and cache block: 2 steps untimed, then trial_steps timed
  auto trial = [&](double* us) -> int {
    b->built = false;
    VI_TRY(batch_prepare(b));
    VI_TRY(batch_launch(b, 2));
    VI_HIP(hipEventRecord(e0, b->stream));
    VI_TRY(batch_launch(b, trial_steps));
    VI_HIP(hipEventRecord(e1, b->stream));
    VI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *us = 1e3 * ms / trial_steps;
    return restore();
  };
  const int cands[4] = {1, 2, 4, 8};
  int best = b->rz;
  double best_us = 1e300;
  int rc = 0;
  for (int c = 0; c < 4 && !rc; ++c) {
    b->rz = cands[c];
    double us = 0;
    if ((rc = trial(&us))) break;
    if (us_per_step) us_per_step[c] = us;
    if (us < best_us) {
      best_us = us;
      best = cands[c];
    }
  }
  // cache blocking over shots against lockstep, at the chosen tiling
  const size_t blk = b->block_bytes;
  if (!rc && blk) {
    b->rz = best;
    b->block_bytes = 0;
    double us = 0;
    rc = trial(&us);
    if (rc || us >= best_us) b->block_bytes = blk;
  }
  (void)hipStreamSynchronize(b->stream);
  if (rc) (void)restore();
  (void)hipStreamSynchronize(b->stream);
  b->prof.on = prof0;
  b->graph_steps = graph0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return g_err.empty() ? fail(rc, "vi_batch_autotune: HIP error") : rc;
  b->rz = best;
  b->built = false;
  invalidate_graph(b);
  if (chosen) *chosen = best;
  return 0;
}

int vi_batch_profile_enable(vi_batch b, int every) {
  if (!b || every < 0) return fail(VI_EINVAL, "vi_batch_profile_enable: bad arguments");
  VI_TRY(b->prof.collect(b->stream));
  b->prof.on = every != 0;
  b->prof.every = every ? every : 1;
  b->prof.reset();
  return 0;
}

int vi_batch_profile_read(vi_batch b, int kernel, double* ms_total, long* launches) {
  if (!b || kernel < 0 || kernel > 1 || !ms_total || !launches)
    return fail(VI_EINVAL, "vi_batch_profile_read: bad arguments");
  VI_TRY(b->prof.collect(b->stream));
  *ms_total = b->prof.ms[kernel];
  *launches = b->prof.n[kernel];
  return 0;
}

}  // extern "C"
