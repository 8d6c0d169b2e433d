// vi_kernels.hpp -- gfx950 kernels of the fused velocity-stress step.
//
// One step = k_velocity (vx, vz from the stress divergence) then k_stress (sxx, szz, sxz
// and the SLS memory variables from the velocity gradients). A launch advances a batch of
// shots: every shot is described by a ShotDesc in device memory, its interior
// [4, nx-4) x [4, nz-4) is cut into tiles of TX x (NW*RZ) cells, and the tiles of all shots
// form one grid (tile_shot[] maps a global tile to its shot). A single simulation is a
// batch of one, launched with a null tile_shot (one dependent load less per workgroup).
//   - a workgroup is NW wavefronts of 64 lanes; lane = x offset, wave = band of RZ rows;
//   - the differentiated fields of the tile are staged once in LDS with a 4-cell halo in x
//     and z (the halo is shared by the NW wavefronts, so each global row segment is a
//     coalesced 64-lane load issued once per workgroup); cell-local fields (state being
//     updated, model, memory variables) stream straight from HBM to registers;
//   - RZ (rows per wave, so NW*RZ rows per tile) is a template parameter chosen by the
//     autotuner: small grids need small tiles for enough waves in flight, large grids
//     amortise the 8-row z halo over taller tiles;
//   - the workgroup -> tile map is XCD-aware: each of the 8 XCDs gets one contiguous range
//     of tiles, so vertical halos are re-read from that XCD's L2.
// State planes are padded (pitch ldx >= nx + 72, rows >= nz + NW*RZ_MAX + 16) and so are
// model planes, so halo and window loads of edge tiles stay inside the allocation without
// bounds checks; only stores are guarded. Padding is zero and never written. Per-lane
// addresses are 32-bit byte offsets from uniform plane pointers.
//
// Arithmetic follows the oracle's operation order exactly (no contraction: the library is
// built with -ffp-contract=off), so FP32 and FP64 results match the CPU oracle bit for
// bit. References: operators src/diff-funcs.jl:3-32; velocity src/visco-kernels.jl:3-20;
// constitutive + memory src/visco-kernels.jl:22-85; VTI stiffnesses
// src/sciml_solver/Models.jl:27-33 in the form of ViscoelasticProblem.jl:21-35; CPML
// src/sciml_solver/CPML.jl:291-328; imaging src/sciml_solver/CPML.jl:372-380; H-AFDA free
// surface src/free-surface-kernels.jl:16-214 with src/common/Stencil.jl:52-189; source
// src/sciml_solver/Source.jl:35-40.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace vi {

constexpr int TX = 64;     // tile width: one wavefront along x
constexpr int NW = 4;      // wavefronts per workgroup
constexpr int RZ_MAX = 8;  // largest rows-per-wave variant (sets the padding)
constexpr int MAXL = 6;

// Everything one shot's kernels need. Lives in device memory; read with scalar loads.
template <typename R>
struct ShotDesc {
  int nx, nz, ldx, ldm;  // state and model row pitch (elements)
  int ntx, tile_base;    // tiles along x; first global tile of this shot
  size_t plane;          // elements per padded state plane (M planes are this far apart)
  long step;             // time-step index of kstep = 0 of the current launch
  R dt, Dx[4], Dz[4];
  R a1[MAXL], a2[MAXL];
  R *vx, *vz, *sxx, *szz, *sxz, *M;
  const R *rho, *lam, *mu, *tau;  // model window (column offset applied), pitch ldm
  // VTI (vti = 1): lam holds C13, mu holds C55 and c11, c33 the diagonal stiffnesses;
  // isotropic (vti = 0): C11 = C33 = lam + 2 mu computed in the kernel
  const R *c11, *c33;
  int vti;
  int T, xlo, xhi, zlo, zhi;
  int fs;          // free surface on row 4: 0 none, 1 stress imaging, 2 H-AFDA, 3 H-AFDA with
                   // the Hermitian formula #3 for d vz/dz on row 5
  // fs >= 2: one-sided d/dz weights [13][8] over rows 4..11 (A0..3 B0..2 C1..3 D0..2), the
  // adjoint's difference table [13][8], then formula #3: beta_m / h (m < 8), alpha
  const R* afda;
  const R *ax, *bx, *kx, *az, *bz, *kz;
  R* psi[8];  // 0..3: [nz][2T] (vx_x, vz_x, sxx_x, sxz_x); 4..7: [2T][ldx] (vx_z, vz_z, sxz_z, szz_z)
  // sources / receivers grouped by tile (CSR over this shot's tiles)
  const int *src_off, *src_cell, *src_kind, *src_row;
  const R* src_w;   // [nsrc][snt]
  const R* src_wc;  // [snt][nsrc]: the same samples, sources in the CSR order of src_off
  int snt, nsrc;
  const int *rec_off, *rec_cell, *rec_comp, *rec_id;
  R* traces;  // [rec_nt][nrec]
  int nrec, rec_nt;
};

// Where a recording launch (REC) stores what the adjoint of its step needs from the forward
// pass, for the interior cells (planes of pitch ldx, pb bytes apart): 0, 1 = fx, fz, the
// velocity half's stress divergence (CPML-stretched, free-surface rows and force sources
// included, before the buoyancy); 2, 3, 4 = exx, ezz, gxz, the stress half's CPML-stretched
// strain rates (ezz before the free-surface replacement); 5..7 = the memory-variable sums of
// xx, zz, xz before the step. The adjoint then needs neither the fields' halos nor the CPML
// memory of the step. Only written by the variants with REC = true.
struct RecSlot {
  char* base;
  unsigned long pb;
};

template <typename R>
__device__ __forceinline__ R* rec_plane(const RecSlot& rs, int q) {
  return (R*)(rs.base + (unsigned long)q * rs.pb);
}

// Workgroup id -> global tile id such that each of the 8 XCDs (which receive workgroups
// round robin) owns one contiguous range of tiles.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7;
  const int x = b & 7, k = b >> 3;
  return x * per + (x < rem ? x : rem) + k;
}

// D+ at the staggered position: sum_l D[l] (u[c+l] - u[c-l+1]), u[k] = q[k*st]
template <typename R>
__device__ __forceinline__ R op_p(const R* q, int st, const R* D) {
  R v = D[0] * (q[st] - q[0]);
  v = v + D[1] * (q[2 * st] - q[-st]);
  v = v + D[2] * (q[3 * st] - q[-2 * st]);
  v = v + D[3] * (q[4 * st] - q[-3 * st]);
  return v;
}
// D-: sum_l D[l] (u[c+l-1] - u[c-l])
template <typename R>
__device__ __forceinline__ R op_m(const R* q, int st, const R* D) {
  R v = D[0] * (q[0] - q[-st]);
  v = v + D[1] * (q[st] - q[-2 * st]);
  v = v + D[2] * (q[2 * st] - q[-3 * st]);
  v = v + D[3] * (q[3 * st] - q[-4 * st]);
  return v;
}

// H-AFDA one-sided d/dz from grid rows 4..11 of a staged tile column (q = row 4, st = LDS
// row pitch) with the weights of table row tr (wave-uniform), accumulated in the same order
// as the oracle: w0 q0, then + w_m q_m. The weights are read through the scalar cache.
template <typename R>
__device__ __forceinline__ R afda_dz(const R* q, int st, const R* tab, int tr) {
  const __attribute__((address_space(4))) R* w =
      (const __attribute__((address_space(4))) R*)tab + __builtin_amdgcn_readfirstlane(tr * 8);
  R v = w[0] * q[0];
#pragma unroll
  for (int m = 1; m < 8; ++m) v = v + w[m] * q[m * st];
  return v;
}

// H-AFDA formula #3 on row 5 (fs = 3): d vz/dz at z = h = alpha es + sum_m beta_m vz(row 4+m),
// es = d vz/dz at the surface, accumulated in the oracle's order (alpha es first). The
// weights sit at tab[208..215] (beta / h) and tab[216] (alpha).
template <typename R>
__device__ __forceinline__ R herm_dz(const R* q, int st, const R* tab, R es) {
  const __attribute__((address_space(4))) R* w = (const __attribute__((address_space(4))) R*)tab;
  R v = w[216] * es;
#pragma unroll
  for (int m = 0; m < 8; ++m) v = v + w[208 + m] * q[m * st];
  return v;
}

// Global-memory access through a uniform base pointer (SGPR pair) plus a 32-bit per-lane
// byte offset: compiles to global_load/store in saddr mode (no 64-bit per-lane address
// arithmetic, no FLAT instructions that would also count against the LDS wait counter).
// Every plane is smaller than 2 GiB (checked when a handle is created).
template <typename T>
__device__ __forceinline__ T gld(const T* base, unsigned boff) {
  return *(const __attribute__((address_space(1))) T*)(
      (const __attribute__((address_space(1))) char*)base + boff);
}
template <typename T>
__device__ __forceinline__ void gst(T* base, unsigned boff, T v) {
  *(__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) char*)base + boff) = v;
}

// The same with the non-temporal hint, for data touched once per step whose next use is far
// away (the per-step slots of the gradient), so that streaming it does not evict the working
// set that the next launches read again from the Infinity Cache.
template <typename T>
__device__ __forceinline__ T gld_nt(const T* base, unsigned boff) {
  return __builtin_nontemporal_load((const __attribute__((address_space(1))) T*)(
      (const __attribute__((address_space(1))) char*)base + boff));
}
template <typename T>
__device__ __forceinline__ void gst_nt(T* base, unsigned boff, T v) {
  __builtin_nontemporal_store(
      v, (__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) char*)base + boff));
}

// Load of a wave-uniform index through the scalar data cache (read-only data): waiting for it
// counts on lgkmcnt only, so it does not drain the vector loads already in flight.
__device__ __forceinline__ int sld(const int* p, int idx) {
  return ((const __attribute__((address_space(4))) int*)p)[__builtin_amdgcn_readfirstlane(idx)];
}

// d / kappa, skipped by the whole wavefront when no lane has kappa != 1 (x / 1 = x exactly, so
// the result is the same bit for bit): the usual kappa_max = 1 strips make no division. The
// test is a wave-uniform (scalar) branch; the empty statement in it keeps the compiler from
// computing the division anyway and selecting (which it does for a plain if).
template <typename R>
__device__ __forceinline__ R div_kappa(R d, R k) {
  if (__builtin_amdgcn_ballot_w64(k != R(1)) != 0) {
    asm volatile("");
    d = d / k;
  }
  return d;
}

// CPML on one derivative d of a strip cell: psi = b psi + a d; return d/kappa + psi.
template <typename R>
__device__ __forceinline__ R cpml_apply(R* psi, unsigned pb, R a, R b, R k, R d) {
  const R q = b * gld(psi, pb) + a * d;
  gst(psi, pb, q);
  return div_kappa(d, k) + q;
}

// The same derivative without storing psi (what the next update would use).
template <typename R>
__device__ __forceinline__ R cpml_rate(const R* psi, unsigned pb, R a, R b, R k, R d) {
  const R q = b * gld(psi, pb) + a * d;
  return div_kappa(d, k) + q;
}

__device__ __forceinline__ int strip_x(int i, int nx, int T, int lo, int hi) {
  if (lo && i < T) return i;
  if (hi && i >= nx - T) return i - (nx - 2 * T);
  return -1;
}

// d vz/dz at the surface (row 4, column i) that keeps szz = 0 there, from the current state
// and without storing anything: the free-surface row's substitute ezz, the vz,z(0) that
// formula #3 takes (with L = 0 and isotropy -lam/(lam + 2 mu) d vx/dx,
// src/free-surface-kernels.jl:160-172). exx = the raw D- d vx/dx of row 4.
template <typename R, int L>
__device__ __forceinline__ R surface_ezz(const ShotDesc<R>& d, int i, R exx) {
  constexpr unsigned E = sizeof(R);
  const int T = d.T;
  if (T) {
    const int xk = strip_x(i, d.nx, T, d.xlo, d.xhi);
    if (xk >= 0) {
      const unsigned q = (unsigned)(4 * 2 * T + xk) * E;
      exx = cpml_rate(d.psi[0], q, gld(d.ax, i * E), gld(d.bx, i * E), gld(d.kx, i * E), exx);
    }
  }
  const unsigned mb = (unsigned)(4 * d.ldm + i) * E, cb = (unsigned)(4 * d.ldx + i) * E;
  const R lr = gld(d.lam, mb);
  const R pz = d.vti ? gld(d.c33, mb) : lr + R(2) * gld(d.mu, mb);
  if constexpr (L > 0) {
    const R t = gld(d.tau, mb);
    const R pz1 = pz * t, pz0 = pz1 + pz;
    const R l1 = lr * t, l0 = l1 + lr;
    R SMxx = gld((const R*)d.M, cb), SMzz = gld((const R*)d.M + (size_t)L * d.plane, cb);
#pragma unroll
    for (int k = 1; k < L; ++k) {
      SMxx = SMxx + gld((const R*)d.M + (size_t)k * d.plane, cb);
      SMzz = SMzz + gld((const R*)d.M + (size_t)(L + k) * d.plane, cb);
    }
    return -(((l0 * exx) + l1 * SMxx) + pz1 * SMzz) / pz0;
  } else {
    return -(lr * exx) / pz;
  }
}

// Cooperative load of NROWS rows (starting at global row jr0) of one field into an LDS
// tile of width TX + 2*HALO (HALO = 4 or 0). Wave w owns rows w, w+NW, ...; lane = main
// column, lanes 0..7 also fetch the 4 + 4 halo columns. fetch() only issues the global
// loads into registers and store() writes them to LDS, so a kernel can put every load of
// all its tiles in flight before the first LDS write (one wait instead of a round trip
// per tile). Rows past NROWS re-load the last row and are not written.
template <typename R, int NROWS, int HALO, int NWV = NW>
struct TileLoad {
  static constexpr int W = TX + 2 * HALO;
  static constexpr int K = (NROWS + NWV - 1) / NWV;
  R v[K], hv[K];
  __device__ __forceinline__ void fetch(const R* __restrict__ g, int jr0, int i0, int ldx,
                                        int lane, int w) {
    constexpr unsigned E = sizeof(R);
    const int hg = lane < 4 ? lane - 4 : TX + lane - 4;  // halo column relative to i0
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int r = min(w + k * NWV, NROWS - 1);
      const unsigned rb = (unsigned)((jr0 + r) * ldx + i0) * E;
      v[k] = gld(g, rb + lane * E);
      if (HALO) hv[k] = lane < 8 ? gld(g, rb + hg * (int)E) : R(0);
    }
  }
  __device__ __forceinline__ void store(R* sh, int lane, int w) const {
    const int hc = lane < 4 ? lane : lane + TX;  // LDS column of a halo value
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int r = w + k * NWV;
      if (r < NROWS) {
        sh[r * W + lane + HALO] = v[k];
        if (HALO && lane < 8) sh[r * W + hc] = hv[k];
      }
    }
  }
};

// FP32 tile rows go straight from global memory to LDS (global_load_lds_dword: lane l of a
// wavefront writes LDS at the wave-uniform base + 4 l), which needs no VGPRs and no LDS stores
// for the 64 main columns; the 4 + 4 halo columns of a row go through registers (lanes 0..7),
// since an LDS-direct load has one base address per instruction and the two halo groups need
// two (a divergent pair of such loads can be merged by the compiler into one with the first
// lane's base, which is wrong). FP64 tiles go through registers (TileLoad): there is no 8-byte
// LDS-direct load. Same rows, columns and LDS layout as TileLoad; the barrier after the loads
// waits for them. Rows past NROWS are not loaded. DMA = false selects the register path for
// FP32 too.
__device__ __forceinline__ void dma_f32(const float* g, float* lds) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) float*)g,
                                   (__attribute__((address_space(3))) float*)lds, 4, 0, 0);
}
template <typename R, int NROWS, int HALO, bool DMA = true, int NWV = NW>
struct TileIn {
  static constexpr int W = TX + 2 * HALO;
  static constexpr int K = (NROWS + NWV - 1) / NWV;
  TileLoad<R, NROWS, HALO, NWV> t;  // FP64 path
  R hv[K];                     // FP32 path: halo values of lanes 0..7
  __device__ __forceinline__ void fetch(const R* __restrict__ g, int jr0, int i0, int ldx,
                                        int lane, int w, R* sh) {
    if constexpr (sizeof(R) == 4 && DMA) {
      constexpr unsigned E = sizeof(R);
      const int hg = lane < 4 ? lane - 4 : TX + lane - 4;  // halo column relative to i0
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int r = w + k * NWV;  // wave-uniform
        if (r < NROWS) dma_f32(g + (size_t)(jr0 + r) * ldx + i0 + lane, &sh[r * W + HALO]);
        if (HALO) {
          const unsigned rb = (unsigned)((jr0 + min(r, NROWS - 1)) * ldx + i0) * E;
          hv[k] = lane < 8 ? gld(g, rb + hg * (int)E) : R(0);
        }
      }
    } else {
      t.fetch(g, jr0, i0, ldx, lane, w);
    }
  }
  __device__ __forceinline__ void store(R* sh, int lane, int w) const {
    if constexpr (sizeof(R) == 4 && DMA) {
      if (HALO) {
        const int hc = lane < 4 ? lane : lane + TX;  // LDS column of a halo value
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int r = w + k * NWV;
          if (r < NROWS && lane < 8) sh[r * W + hc] = hv[k];
        }
      }
    } else {
      t.store(sh, lane, w);
    }
  }
};

// CPML memory of one lane's RZ cells for two x-strip and two z-strip arrays, requested
// before the barrier together with the tile loads (for short tiles), so the strip tiles,
// which finish last, do not wait for a separate round trip after the stencils.
template <typename R, int RZ>
struct CpmlRows {
  R px0[RZ], px1[RZ], pz0[RZ], pz1[RZ];
  R ax = 0, bx = 0, kx = 1;
  int xk = -1;
  __device__ __forceinline__ void fetch(const ShotDesc<R>& d, int i, int jw, bool col_ok,
                                        int ix0, int ix1, int iz0, int iz1) {
    constexpr unsigned E = sizeof(R);
    const int T = d.T;
    xk = (T && col_ok) ? strip_x(i, d.nx, T, d.xlo, d.xhi) : -1;
    if (xk >= 0) {
      ax = gld(d.ax, i * E);
      bx = gld(d.bx, i * E);
      kx = gld(d.kx, i * E);
    }
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int j = jw + r;
      const bool ok = col_ok && j < d.nz - 4;
      px0[r] = px1[r] = pz0[r] = pz1[r] = R(0);
      if (ok && xk >= 0) {
        const unsigned q = (unsigned)(j * 2 * T + xk) * E;
        px0[r] = gld((const R*)d.psi[ix0], q);
        px1[r] = gld((const R*)d.psi[ix1], q);
      }
      const int zk = (T && ok) ? strip_x(j, d.nz, T, d.zlo, d.zhi) : -1;
      if (zk >= 0) {
        const unsigned q = (unsigned)(zk * d.ldx + i) * E;
        pz0[r] = gld((const R*)d.psi[iz0], q);
        pz1[r] = gld((const R*)d.psi[iz1], q);
      }
    }
  }
};

// psi' = b psi + a d (stored), return d/kappa + psi' -- with psi already in a register
template <typename R>
__device__ __forceinline__ R cpml_apply_pre(R* psi, unsigned pb, R p, R a, R b, R k, R d) {
  const R q = b * p + a * d;
  gst(psi, pb, q);
  return div_kappa(d, k) + q;
}

// The first 64 entries of a tile's source and receiver lists (CSR order), copied to LDS by
// the first wavefront together with the tile loads, before the tile's barrier. The per-cell
// loops then read them at uniform LDS addresses instead of making one dependent global
// round trip per entry (which made the tiles with many receivers the slowest of a launch).
// Entries past 64 are read from global memory.
constexpr int LIST = 64;
template <typename R>
struct TileLists {
  int src_cell[LIST], src_kind[LIST];
  R src_w[LIST];  // the source's sample of this step
  int rec_cell[LIST], rec_id[LIST], rec_comp[LIST];
};

// The list entries, requested by the first wavefront together with the tile loads; no load
// depends on another one (the samples are stored per step in CSR order). FP32: straight into
// LDS like the tiles (no VGPRs); FP64: through registers, written to LDS with the tiles.
__device__ __forceinline__ void dma4(const void* g, void* lds) {
  dma_f32((const float*)g, (float*)lds);
}
template <typename R>
struct ListIn {
  int scell = 0, skind = 0, rcell = 0, rid = 0, rcomp = 0;  // FP64 path only
  R sw = R(0);
  __device__ __forceinline__ void fetch(TileLists<R>& tl, const ShotDesc<R>& d, int s0, int s1,
                                        int r0, int r1, long step, int lane, int w,
                                        bool with_receivers) {
    constexpr unsigned E = sizeof(R);
    if (w != 0) return;
    if (lane < s1 - s0) {
      const unsigned s = s0 + lane;
      if constexpr (sizeof(R) == 4) {
        dma4(d.src_cell + s, tl.src_cell);
        dma4(d.src_kind + s, tl.src_kind);
        if (step < d.snt) dma4(d.src_wc + (size_t)step * d.nsrc + s, tl.src_w);
        else tl.src_w[lane] = R(0);
      } else {
        scell = gld(d.src_cell, s * 4u);
        skind = gld(d.src_kind, s * 4u);
        if (step < d.snt) sw = gld(d.src_wc, ((unsigned)step * d.nsrc + s) * E);
      }
    }
    if (with_receivers && lane < r1 - r0) {
      const unsigned r = r0 + lane;
      if constexpr (sizeof(R) == 4) {
        dma4(d.rec_cell + r, tl.rec_cell);
        dma4(d.rec_id + r, tl.rec_id);
        dma4(d.rec_comp + r, tl.rec_comp);
      } else {
        rcell = gld(d.rec_cell, r * 4u);
        rid = gld(d.rec_id, r * 4u);
        rcomp = gld(d.rec_comp, r * 4u);
      }
    }
  }
  __device__ __forceinline__ void store(TileLists<R>& tl, int s0, int s1, int r0, int r1,
                                        int lane, int w, bool with_receivers) const {
    if constexpr (sizeof(R) != 4) {
      if (w != 0) return;
      if (lane < s1 - s0) {
        tl.src_cell[lane] = scell;
        tl.src_kind[lane] = skind;
        tl.src_w[lane] = sw;
      }
      if (with_receivers && lane < r1 - r0) {
        tl.rec_cell[lane] = rcell;
        tl.rec_id[lane] = rid;
        tl.rec_comp[lane] = rcomp;
      }
    }
  }
};

// AFDA: the variant with the H-AFDA rows (any shot of the launch may use them; the others
// run the same code as without)
template <typename R, int RZ, bool REC = false, bool AFDA = false, int NWV = NW>
__global__ __launch_bounds__(TX* NWV) void k_velocity(const ShotDesc<R>* __restrict__ shots,
                                                     const int* __restrict__ tile_shot,
                                                     int kstep, const RecSlot rs) {
  static_assert(!REC || RZ <= 2, "recording variants exist for 1 and 2 rows per wave");
  constexpr int TZ = NWV * RZ;
  constexpr int WX = TX + 8;
  constexpr unsigned E = sizeof(R);
  __shared__ R s_xz[(TZ + 8) * WX];  // sxz rows j0-4 .. j0+TZ+3, columns i0-4 .. i0+67
  __shared__ R s_zz[(TZ + 8) * TX];  // szz rows j0-4 .. j0+TZ+3, columns i0 .. i0+63
  __shared__ R s_xx[TZ * WX];        // sxx rows j0 .. j0+TZ-1, columns i0-4 .. i0+67
  const int lane = threadIdx.x, w = threadIdx.y;
  const int gtile = xcd_tile(blockIdx.x, gridDim.x);
  const ShotDesc<R>& d = shots[tile_shot ? tile_shot[gtile] : 0];  // null: one shot
  const int tile = gtile - d.tile_base;
  const int ntx = d.ntx;
  const int tx = tile % ntx, tz = tile / ntx;
  const int i0 = 4 + tx * TX, j0 = 4 + tz * TZ;
  const int i = i0 + lane;
  const int ldx = d.ldx, ldm = d.ldm;
  // register path (DMA = false): the LDS-direct rows measured 0.4 us slower in this kernel
  // at 1024x512 (same-box A/B), while they save 0.2-0.8 us in the stress kernel
  TileIn<R, TZ + 8, 4, false, NWV> t_xz;
  TileIn<R, TZ + 8, 0, false, NWV> t_zz;
  TileIn<R, TZ, 4, false, NWV> t_xx;
  t_xz.fetch(d.sxz, j0 - 4, i0, ldx, lane, w, s_xz);
  t_zz.fetch(d.szz, j0 - 4, i0, ldx, lane, w, s_zz);
  t_xx.fetch(d.sxx, j0, i0, ldx, lane, w, s_xx);

  const int jw = j0 + w * RZ;  // first row of this wave
  const unsigned rowb = (unsigned)ldx * E, mrowb = (unsigned)ldm * E;
  const unsigned cb0 = (unsigned)(jw * ldx + i) * E;   // byte offset of (i, jw) in a state plane
  const unsigned mb0 = (unsigned)(jw * ldm + i) * E;   // ... and in a model plane
  const R* __restrict__ rho = d.rho;
  R* __restrict__ vx = d.vx;
  R* __restrict__ vz = d.vz;
  constexpr bool PRE = RZ <= 2;  // CPML memory requested before the barrier
  CpmlRows<R, PRE ? RZ : 1> cp;
  if constexpr (PRE) cp.fetch(d, i, jw, i < d.nx - 4, 2, 3, 6, 7);
  R wrho[RZ + 1], wrho1[RZ], ovx[RZ], ovz[RZ];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    wrho1[r] = gld(rho, mb0 + r * mrowb + E);
    ovx[r] = gld((const R*)vx, cb0 + r * rowb);
    ovz[r] = gld((const R*)vz, cb0 + r * rowb);
  }
#pragma unroll
  for (int r = 0; r <= RZ; ++r) wrho[r] = gld(rho, mb0 + r * mrowb);
  const int s0 = sld(d.src_off, tile), s1 = sld(d.src_off, tile + 1);
  const long step = d.step + kstep;
  __shared__ TileLists<R> tl;
  ListIn<R> lr;
  lr.fetch(tl, d, s0, s1, 0, 0, step, lane, w, false);
  t_xz.store(s_xz, lane, w);
  t_zz.store(s_zz, lane, w);
  t_xx.store(s_xx, lane, w);
  lr.store(tl, s0, s1, 0, 0, lane, w, false);
  __syncthreads();

  const int nx = d.nx, nz = d.nz, T = d.T;
  const bool col_ok = i < nx - 4;
  const bool afda = AFDA && d.fs >= 2;
  const R dt = d.dt, two = R(2);
  const R Dx[4] = {d.Dx[0], d.Dx[1], d.Dx[2], d.Dx[3]};
  const R Dz[4] = {d.Dz[0], d.Dz[1], d.Dz[2], d.Dz[3]};
  const int xk = T ? strip_x(i, nx, T, d.xlo, d.xhi) : -1;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = jw + r;
    const int tr = w * RZ + r;  // row within the tile
    const unsigned cb = cb0 + r * rowb;
    R dsxx_x = op_p(&s_xx[tr * WX + lane + 4], 1, Dx);
    R dsxz_x = op_m(&s_xz[(tr + 4) * WX + lane + 4], 1, Dx);
    R dsxz_z = op_m(&s_xz[(tr + 4) * WX + lane + 4], WX, Dz);
    R dszz_z = op_p(&s_zz[(tr + 4) * TX + lane], TX, Dz);
    // H-AFDA rows (only in the first tile row, where LDS row = grid row); ju is j, known to
    // the compiler as wave-uniform, so this is a scalar branch
    const int ju = __builtin_amdgcn_readfirstlane(jw) + r;
    if (AFDA && afda && ju < 8) {
      dsxz_z = afda_dz(&s_xz[4 * WX + lane + 4], WX, d.afda, ju - 4);      // A_{j-4}
      if (ju < 7) dszz_z = afda_dz(&s_zz[4 * TX + lane], TX, d.afda, ju);  // B_{j-4}
    }
    const bool ok = col_ok && j < nz - 4;
    if (T && ok) {
      if (xk >= 0) {
        const unsigned q = (unsigned)(j * 2 * T + xk) * E;
        if constexpr (PRE) {
          dsxx_x = cpml_apply_pre(d.psi[2], q, cp.px0[r], cp.ax, cp.bx, cp.kx, dsxx_x);
          dsxz_x = cpml_apply_pre(d.psi[3], q, cp.px1[r], cp.ax, cp.bx, cp.kx, dsxz_x);
        } else {
          const R a = gld(d.ax, i * E), b = gld(d.bx, i * E), k = gld(d.kx, i * E);
          dsxx_x = cpml_apply(d.psi[2], q, a, b, k, dsxx_x);
          dsxz_x = cpml_apply(d.psi[3], q, a, b, k, dsxz_x);
        }
      }
      const int zk = strip_x(j, nz, T, d.zlo, d.zhi);
      if (zk >= 0) {
        const unsigned q = (unsigned)(zk * ldx + i) * E;
        const R a = d.az[j], b = d.bz[j], k = d.kz[j];
        if constexpr (PRE) {
          dsxz_z = cpml_apply_pre(d.psi[6], q, cp.pz0[r], a, b, k, dsxz_z);
          dszz_z = cpml_apply_pre(d.psi[7], q, cp.pz1[r], a, b, k, dszz_z);
        } else {
          dsxz_z = cpml_apply(d.psi[6], q, a, b, k, dsxz_z);
          dszz_z = cpml_apply(d.psi[7], q, a, b, k, dszz_z);
        }
      }
    }
    R fx = dsxx_x + dsxz_z;
    R fz = dszz_z + dsxz_x;
    const unsigned cell = cb / E;
    for (int q = 0; q < min(s1 - s0, LIST); ++q) {  // sources in CSR order
      if ((unsigned)tl.src_cell[q] != cell) continue;
      const int kind = tl.src_kind[q];
      const R val = tl.src_w[q];
      if (kind == 0) fz = fz + val;
      else if (kind == 2) fx = fx + val;
    }
    for (int s = s0 + LIST; s < s1; ++s) {
      if ((unsigned)d.src_cell[s] * E != cb) continue;
      const int kind = d.src_kind[s];
      const R val = step < d.snt ? d.src_wc[(size_t)step * d.nsrc + s] : R(0);
      if (kind == 0) fz = fz + val;
      else if (kind == 2) fx = fx + val;
    }
    const R bxv = two / (wrho[r] + wrho1[r]);
    const R bzv = two / (wrho[r] + wrho[r + 1]);
    if (ok) {
      const R nvx = ovx[r] + dt * (bxv * fx), nvz = ovz[r] + dt * (bzv * fz);
      gst(vx, cb, nvx);
      gst(vz, cb, nvz);
      if constexpr (REC) {
        gst_nt(rec_plane<R>(rs, 0), cb, fx);
        gst_nt(rec_plane<R>(rs, 1), cb, fz);
      }
    }
  }
}

template <typename R, int L, int RZ, bool REC, bool AFDA>
constexpr int stress_waves() {
  return (sizeof(R) == 4 && RZ == 1 && L <= 3 && !REC && !AFDA) ? 8 : 1;
}

template <typename R, int L, int RZ, bool REC = false, bool AFDA = false, int NWV = NW>
__global__ __launch_bounds__(TX* NWV, (stress_waves<R, L, RZ, REC, AFDA>())) void k_stress(const ShotDesc<R>* __restrict__ shots,
                                                   const int* __restrict__ tile_shot,
                                                   int kstep, const RecSlot rs) {
  static_assert(!REC || RZ <= 2, "recording variants exist for 1 and 2 rows per wave");
  constexpr int TZ = NWV * RZ;
  constexpr int WX = TX + 8;
  constexpr unsigned E = sizeof(R);
  // Cell-local operands of all RZ rows are requested before the barrier (in flight with
  // the tile loads) for short tiles; taller tiles load them row by row to keep VGPRs low.
  constexpr bool HOIST = RZ <= 2;
  constexpr int LM = L > 0 ? L : 1;
  constexpr int RH = HOIST ? RZ : 1;
  __shared__ R s_vx[(TZ + 8) * WX];  // vx rows j0-4 .. j0+TZ+3, columns i0-4 .. i0+67
  __shared__ R s_vz[(TZ + 8) * WX];  // vz, same window
  __shared__ R s_es[AFDA ? TX : 1];  // formula #3: d vz/dz at the surface per column
  const int lane = threadIdx.x, w = threadIdx.y;
  const int gtile = xcd_tile(blockIdx.x, gridDim.x);
  const ShotDesc<R>& d = shots[tile_shot ? tile_shot[gtile] : 0];  // null: one shot
  const int tile = gtile - d.tile_base;
  const int ntx = d.ntx;
  const int tx = tile % ntx, tz = tile / ntx;
  const int i0 = 4 + tx * TX, j0 = 4 + tz * TZ;
  const int i = i0 + lane;
  const int ldx = d.ldx, ldm = d.ldm;
  TileIn<R, TZ + 8, 4, true, NWV> t_vx, t_vz;
  t_vx.fetch(d.vx, j0 - 4, i0, ldx, lane, w, s_vx);
  t_vz.fetch(d.vz, j0 - 4, i0, ldx, lane, w, s_vz);

  const int jw = j0 + w * RZ;
  const unsigned rowb = (unsigned)ldx * E, mrowb = (unsigned)ldm * E;
  const unsigned cb0 = (unsigned)(jw * ldx + i) * E;
  const unsigned mb0 = (unsigned)(jw * ldm + i) * E;
  const R* __restrict__ mu = d.mu;
  const R* __restrict__ lam = d.lam;
  const R* __restrict__ tau = d.tau;
  R* __restrict__ sxx = d.sxx;
  R* __restrict__ szz = d.szz;
  R* __restrict__ sxz = d.sxz;
  R* Mp[3 * LM];
#pragma unroll
  for (int k = 0; k < 3 * LM; ++k) Mp[k] = d.M + (size_t)k * d.plane;
  R wmu[RZ + 1], wmu1[RZ + 1];  // mu at columns i and i+1, rows jw .. jw+RZ
  // CPML memory before the barrier: pays for 2-row tiles (measured); the 1-row variant
  // keeps its registers for occupancy
  constexpr bool PRE = RZ == 2;
  CpmlRows<R, PRE ? RZ : 1> cp;
  if constexpr (PRE) cp.fetch(d, i, jw, i < d.nx - 4, 0, 1, 4, 5);
  const long step = d.step + kstep;
  __shared__ TileLists<R> tl;
  const bool vti = d.vti;
  // per cell: lam | C13, C11, C33 (lam + 2 mu when isotropic), mu averaged to the sxz node,
  // tau, the stress and the memory variables; the moduli are formed as soon as the loads land,
  // so the four mu values are not held through the CPML code (fewer VGPRs)
  R olam[RH], otau[RH], osxx[RH], oszz[RH], osxz[RH], om[3 * LM][RH], oc11[RH], oc33[RH];
  R omxz[RH];
  const R two = R(2), quarter = R(0.25);
  auto load_cell = [&](int r, int h) {
    const unsigned cb = cb0 + r * rowb, mb = mb0 + r * mrowb;
    olam[h] = gld(lam, mb);
    if (vti) {
      oc11[h] = gld(d.c11, mb);
      oc33[h] = gld(d.c33, mb);
    }
    osxx[h] = gld((const R*)sxx, cb);
    oszz[h] = gld((const R*)szz, cb);
    osxz[h] = gld((const R*)sxz, cb);
    if constexpr (L > 0) {
      otau[h] = gld(tau, mb);
#pragma unroll
      for (int k = 0; k < 3 * L; ++k) om[k][h] = gld((const R*)Mp[k], cb);
    } else {
      otau[h] = R(0);
    }
    if (!vti) {
      oc11[h] = olam[h] + two * wmu[r];  // C11 = C33 = lam + 2 mu
      oc33[h] = oc11[h];
    }
    omxz[h] = quarter * (((wmu[r] + wmu1[r]) + wmu[r + 1]) + wmu1[r + 1]);
    // keep the compiler from sinking these sums to their uses, which would hold the four mu
    // values through the CPML code (an empty statement that only pins the value in a VGPR)
    asm volatile("" : "+v"(omxz[h]), "+v"(oc11[h]), "+v"(oc33[h]));
  };
#pragma unroll
  for (int r = 0; r <= RZ; ++r) {
    wmu[r] = gld(mu, mb0 + r * mrowb);
    wmu1[r] = gld(mu, mb0 + r * mrowb + E);
  }
  if constexpr (HOIST) {
#pragma unroll
    for (int r = 0; r < RZ; ++r) load_cell(r, r);
  }
  const int s0 = sld(d.src_off, tile), s1 = sld(d.src_off, tile + 1);
  const int r0 = sld(d.rec_off, tile), r1 = sld(d.rec_off, tile + 1);
  ListIn<R> lr;
  lr.fetch(tl, d, s0, s1, r0, r1, step, lane, w, true);
  t_vx.store(s_vx, lane, w);
  t_vz.store(s_vz, lane, w);
  lr.store(tl, s0, s1, r0, r1, lane, w, true);
  __syncthreads();

  const int nx = d.nx, nz = d.nz, T = d.T;
  const bool col_ok = i < nx - 4;
  const R dt = d.dt;
  const R Dx[4] = {d.Dx[0], d.Dx[1], d.Dx[2], d.Dx[3]};
  const R Dz[4] = {d.Dz[0], d.Dz[1], d.Dz[2], d.Dz[3]};
  const int xk = T ? strip_x(i, nx, T, d.xlo, d.xhi) : -1;
  const int fs = d.fs;
  if constexpr (AFDA) {
    // formula #3: row 5 needs row 4's surface value, taken here before any wave of the tile
    // updates row 4's psi and memory variables (fs and tz are uniform over the block)
    if (fs == 3 && tz == 0) {
      if (w == 0 && col_ok)
        s_es[lane] = surface_ezz<R, L>(d, i, op_m(&s_vx[4 * WX + lane + 4], 1, Dx));
      __syncthreads();
    }
  }
  R a1[LM], a2[LM];
#pragma unroll
  for (int k = 0; k < L; ++k) {
    a1[k] = d.a1[k];
    a2[k] = d.a2[k];
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = jw + r;
    const int tr = w * RZ + r + 4;  // LDS row of this output
    const unsigned cb = cb0 + r * rowb;
    const int h = HOIST ? r : 0;
    const R* pvx = &s_vx[tr * WX + lane + 4];
    const R* pvz = &s_vz[tr * WX + lane + 4];
    R exx = op_m(pvx, 1, Dx);
    R exz = op_p(pvz, 1, Dx);
    R ezz = op_m(pvz, WX, Dz);
    R ezx = op_p(pvx, WX, Dz);
    // H-AFDA rows (only in the first tile row, where LDS row = grid row); ju is j, known to
    // the compiler as wave-uniform, so this is a scalar branch
    const int ju = __builtin_amdgcn_readfirstlane(jw) + r;
    if (AFDA && fs >= 2 && ju < 8) {
      if (fs == 3 && ju == 5)  // formula #3
        ezz = herm_dz(&s_vz[4 * WX + lane + 4], WX, d.afda, s_es[lane]);
      else if (ju >= 5)
        ezz = afda_dz(&s_vz[4 * WX + lane + 4], WX, d.afda, ju + 2);  // C_{j-4}
      if (ju < 7) ezx = afda_dz(&s_vx[4 * WX + lane + 4], WX, d.afda, ju + 6);   // D_{j-4}
    }
    const bool ok = col_ok && j < nz - 4;
    if (!ok) continue;
    if constexpr (!HOIST) load_cell(r, 0);
    if (T) {
      if (xk >= 0) {
        const unsigned q = (unsigned)(j * 2 * T + xk) * E;
        if constexpr (PRE) {
          exx = cpml_apply_pre(d.psi[0], q, cp.px0[r], cp.ax, cp.bx, cp.kx, exx);
          exz = cpml_apply_pre(d.psi[1], q, cp.px1[r], cp.ax, cp.bx, cp.kx, exz);
        } else {
          const R a = gld(d.ax, i * E), b = gld(d.bx, i * E), k = gld(d.kx, i * E);
          exx = cpml_apply(d.psi[0], q, a, b, k, exx);
          exz = cpml_apply(d.psi[1], q, a, b, k, exz);
        }
      }
      const int zk = strip_x(j, nz, T, d.zlo, d.zhi);
      if (zk >= 0) {
        const unsigned q = (unsigned)(zk * ldx + i) * E;
        const R a = d.az[j], b = d.bz[j], k = d.kz[j];
        if constexpr (PRE) {
          ezx = cpml_apply_pre(d.psi[4], q, cp.pz0[r], a, b, k, ezx);
          ezz = cpml_apply_pre(d.psi[5], q, cp.pz1[r], a, b, k, ezz);
        } else {
          ezx = cpml_apply(d.psi[4], q, a, b, k, ezx);
          ezz = cpml_apply(d.psi[5], q, a, b, k, ezz);
        }
      }
    }
    const R gxz = ezx + exz;
    if constexpr (REC) {
      gst_nt(rec_plane<R>(rs, 2), cb, exx);
      gst_nt(rec_plane<R>(rs, 3), cb, ezz);
      gst_nt(rec_plane<R>(rs, 4), cb, gxz);
    }
    const R lr = olam[h];
    const R mxz = omxz[h];
    const R px = oc11[h];  // C11
    const R pz = oc33[h];  // C33
    const bool fs_row = fs && j == 4;
    R dsxx, dszz, dsxz;
    if constexpr (L > 0) {
      const R t = otau[h];
      const R px1 = px * t, px0 = px1 + px;
      const R pz1 = pz * t, pz0 = pz1 + pz;
      const R l1 = lr * t, l0 = l1 + lr;
      const R m1 = mxz * t, m0 = m1 + mxz;
      R SMxx = om[0][h], SMzz = om[L][h], SMxz = om[2 * L][h];
#pragma unroll
      for (int k = 1; k < L; ++k) {
        SMxx = SMxx + om[k][h];
        SMzz = SMzz + om[L + k][h];
        SMxz = SMxz + om[2 * L + k][h];
      }
      if constexpr (REC) {
        gst_nt(rec_plane<R>(rs, 5), cb, SMxx);
        gst_nt(rec_plane<R>(rs, 6), cb, SMzz);
        gst_nt(rec_plane<R>(rs, 7), cb, SMxz);
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
    // memory variables first: their old values are dead after this, before the source loop
    if constexpr (L > 0) {
#pragma unroll
      for (int k = 0; k < L; ++k) {
        gst(Mp[k], cb, om[k][h] + dt * (a1[k] * exx + a2[k] * om[k][h]));
        gst(Mp[L + k], cb, om[L + k][h] + dt * (a1[k] * ezz + a2[k] * om[L + k][h]));
        gst(Mp[2 * L + k], cb, om[2 * L + k][h] + dt * (a1[k] * gxz + a2[k] * om[2 * L + k][h]));
      }
    }
    const unsigned cell = cb / E;
    for (int q = 0; q < min(s1 - s0, LIST); ++q) {  // explosive sources in CSR order
      if ((unsigned)tl.src_cell[q] != cell || tl.src_kind[q] != 1) continue;
      const R val = tl.src_w[q];
      dsxx = dsxx + val;
      dszz = dszz + val;
    }
    for (int s = s0 + LIST; s < s1; ++s) {
      if ((unsigned)d.src_cell[s] * E != cb || d.src_kind[s] != 1) continue;
      const R val = step < d.snt ? d.src_wc[(size_t)step * d.nsrc + s] : R(0);
      dsxx = dsxx + val;
      dszz = dszz + val;
    }
    const R nsxx = osxx[h] + dt * dsxx;
    R nszz = oszz[h] + dt * dszz;
    const R nsxz = osxz[h] + dt * dsxz;
    if (fs_row) nszz = R(0);
    gst(sxx, cb, nsxx);
    gst(szz, cb, nszz);
    gst(sxz, cb, nsxz);
    if (fs == 1) {  // stress images above the surface row 4
      if (j >= 5 && j <= 8) gst(szz, (unsigned)((8 - j) * ldx + i) * E, -nszz);
      if (j >= 4 && j <= 7) gst(sxz, (unsigned)((7 - j) * ldx + i) * E, -nsxz);
    }
    if (step < d.rec_nt) {
      for (int q = 0; q < min(r1 - r0, LIST); ++q) {  // receivers of the tile, from LDS
        if ((unsigned)tl.rec_cell[q] != cell) continue;
        const int comp = tl.rec_comp[q];
        const R val = comp == 0   ? pvx[0]
                      : comp == 1 ? pvz[0]
                      : comp == 2 ? nsxx
                      : comp == 3 ? nszz
                                  : nsxz;
        d.traces[(size_t)step * d.nrec + tl.rec_id[q]] = val;
      }
    }
    for (int q = r0 + LIST; q < r1; ++q) {
      if ((unsigned)d.rec_cell[q] * E != cb || step >= d.rec_nt) continue;
      const int comp = d.rec_comp[q];
      const R val = comp == 0   ? pvx[0]
                    : comp == 1 ? pvz[0]
                    : comp == 2 ? nsxx
                    : comp == 3 ? nszz
                                : nsxz;
      d.traces[(size_t)step * d.nrec + d.rec_id[q]] = val;
    }
  }
}

// Set the step index of shot s (stream-ordered, so later launches see it).
template <typename R>
__global__ void k_set_step(ShotDesc<R>* shots, int s, long step) {
  if (blockIdx.x * blockDim.x + threadIdx.x == 0) shots[s].step = step;
}

// Advance the step index of every shot of a batch after a block of k steps.
template <typename R>
__global__ void k_advance(ShotDesc<R>* shots, int nshots, int k) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nshots) shots[s].step += k;
}

}  // namespace vi
