// vi_kernels.hpp -- gfx950 kernels of the fused velocity-stress step.
//
// One step = k_velocity (vx, vz from the stress divergence) then k_stress (sxx, szz, sxz
// and the SLS memory variables from the velocity gradients). A launch advances a batch of
// shots: every shot is described by a ShotDesc in device memory, its interior
// [4, nx-4) x [4, nz-4) is cut into tiles of TX x (NW*RZ) cells, and the tiles of all shots
// form one grid (tile_shot[] maps a global tile to its shot). A single simulation is a
// batch of one.
//   - a workgroup is NW wavefronts of 64 lanes; lane = x offset, wave = band of RZ rows;
//   - fields differentiated along x are staged in LDS with a 4-column halo;
//   - fields differentiated along z are held as a register window of RZ+7 rows per lane
//     (z register blocking), so every global load is a coalesced 64-lane row segment;
//   - RZ (rows per wave) is a template parameter chosen by the autotuner: small grids need
//     small RZ for enough waves in flight, large grids amortise the z halo;
//   - the workgroup -> tile map is XCD-aware: each of the 8 XCDs gets one contiguous range
//     of tiles, so vertical halos are re-read from that XCD's L2.
// State planes are padded (pitch ldx >= nx + 72, rows >= nz + NW*RZ_MAX + 16) and so are
// model planes, so halo and window loads of edge tiles stay inside the allocation without
// bounds checks; only stores are guarded. Padding is zero and never written.
//
// Arithmetic follows the oracle's operation order exactly (no contraction: the library is
// built with -ffp-contract=off), so FP32 and FP64 results match the CPU oracle bit for
// bit. References: operators src/diff-funcs.jl:3-32; velocity src/visco-kernels.jl:3-20;
// constitutive + memory src/visco-kernels.jl:22-85; CPML src/sciml_solver/CPML.jl:291-328;
// imaging src/sciml_solver/CPML.jl:372-380; source src/sciml_solver/Source.jl:35-40.
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
  int T, xlo, xhi, zlo, zhi, fs;
  const R *ax, *bx, *kx, *az, *bz, *kz;
  R* psi[8];  // 0..3: [nz][2T] (vx_x, vz_x, sxx_x, sxz_x); 4..7: [2T][ldx] (vx_z, vz_z, sxz_z, szz_z)
  // sources / receivers grouped by tile (CSR over this shot's tiles)
  const int *src_off, *src_cell, *src_kind, *src_row;
  const R* src_w;  // [nsrc][snt]
  int snt;
  const int *rec_off, *rec_cell, *rec_comp, *rec_id;
  R* traces;  // [rec_nt][nrec]
  int nrec, rec_nt;
};

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

template <typename R>
__device__ __forceinline__ R cpml_apply(R* psi, R a, R b, R k, R d) {
  const R q = b * (*psi) + a * d;
  *psi = q;
  return d / k + q;
}

__device__ __forceinline__ int strip_x(int i, int nx, int T, int lo, int hi) {
  if (lo && i < T) return i;
  if (hi && i >= nx - T) return i - (nx - 2 * T);
  return -1;
}

// Stage RZ rows of one field (own column from the register window starting at OFF, halo
// columns from global memory) into this wave's LDS rows.
template <typename R, int RZ, int OFF, int N>
__device__ __forceinline__ void stage_rows(R (*sh)[TX + 8], const R (&win)[N],
                                           const R* __restrict__ g, size_t base, int ldx,
                                           int lane) {
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    sh[r][lane + 4] = win[OFF + r];
    if (lane < 4) sh[r][lane] = g[base + (size_t)r * ldx - 4];
    if (lane >= TX - 4) sh[r][lane + 8] = g[base + (size_t)r * ldx + 4];
  }
}

template <typename R, int RZ>
__global__ __launch_bounds__(TX* NW) void k_velocity(const ShotDesc<R>* __restrict__ shots,
                                                     const int* __restrict__ tile_shot,
                                                     int kstep) {
  __shared__ R sh_xx[NW][RZ][TX + 8];
  __shared__ R sh_xz[NW][RZ][TX + 8];
  const int lane = threadIdx.x, w = threadIdx.y;
  const int gtile = xcd_tile(blockIdx.x, gridDim.x);
  const ShotDesc<R>& d = shots[tile_shot[gtile]];
  const int tile = gtile - d.tile_base;
  const int ntx = d.ntx;
  const int tx = tile % ntx, tz = tile / ntx;
  const int i = 4 + tx * TX + lane;
  const int j0 = 4 + (tz * NW + w) * RZ;
  const int ldx = d.ldx, ldm = d.ldm;
  const size_t base = (size_t)j0 * ldx + i;
  const size_t mbase = (size_t)j0 * ldm + i;
  const R* __restrict__ sxx = d.sxx;
  const R* __restrict__ sxz = d.sxz;
  const R* __restrict__ szz = d.szz;
  const R* __restrict__ rho = d.rho;
  R* __restrict__ vx = d.vx;
  R* __restrict__ vz = d.vz;

  R wxx[RZ];      // sxx rows j0 .. j0+RZ-1
  R wxz[RZ + 7];  // sxz rows j0-4 .. j0+RZ+2   (D- along z, centre at r+4)
  R wzz[RZ + 7];  // szz rows j0-3 .. j0+RZ+3   (D+ along z, centre at r+3)
  R wrho[RZ + 1], wrho1[RZ];  // rho at column i rows j0 .. j0+RZ; column i+1 rows j0 ..
  R ovx[RZ], ovz[RZ];
#pragma unroll
  for (int k = 0; k < RZ + 7; ++k) {
    wxz[k] = sxz[base + (ptrdiff_t)(k - 4) * ldx];
    wzz[k] = szz[base + (ptrdiff_t)(k - 3) * ldx];
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    wxx[r] = sxx[base + (size_t)r * ldx];
    wrho1[r] = rho[mbase + (size_t)r * ldm + 1];
    ovx[r] = vx[base + (size_t)r * ldx];
    ovz[r] = vz[base + (size_t)r * ldx];
  }
#pragma unroll
  for (int r = 0; r <= RZ; ++r) wrho[r] = rho[mbase + (size_t)r * ldm];

  stage_rows<R, RZ, 0>(sh_xx[w], wxx, sxx, base, ldx, lane);
  stage_rows<R, RZ, 4>(sh_xz[w], wxz, sxz, base, ldx, lane);
  __syncthreads();

  const int nx = d.nx, nz = d.nz, T = d.T;
  const bool col_ok = i < nx - 4;
  const R dt = d.dt, two = R(2);
  const R Dx[4] = {d.Dx[0], d.Dx[1], d.Dx[2], d.Dx[3]};
  const R Dz0 = d.Dz[0], Dz1 = d.Dz[1], Dz2 = d.Dz[2], Dz3 = d.Dz[3];
  const int xk = T ? strip_x(i, nx, T, d.xlo, d.xhi) : -1;
  const int s0 = d.src_off[tile], s1 = d.src_off[tile + 1];
  const long step = d.step + kstep;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + r;
    const size_t c = base + (size_t)r * ldx;
    R dsxx_x = op_p(&sh_xx[w][r][lane + 4], 1, Dx);
    R dsxz_x = op_m(&sh_xz[w][r][lane + 4], 1, Dx);
    // D- of sxz along z (centre r+4) and D+ of szz along z (centre r+3)
    R dsxz_z = Dz0 * (wxz[r + 4] - wxz[r + 3]);
    dsxz_z = dsxz_z + Dz1 * (wxz[r + 5] - wxz[r + 2]);
    dsxz_z = dsxz_z + Dz2 * (wxz[r + 6] - wxz[r + 1]);
    dsxz_z = dsxz_z + Dz3 * (wxz[r + 7] - wxz[r + 0]);
    R dszz_z = Dz0 * (wzz[r + 4] - wzz[r + 3]);
    dszz_z = dszz_z + Dz1 * (wzz[r + 5] - wzz[r + 2]);
    dszz_z = dszz_z + Dz2 * (wzz[r + 6] - wzz[r + 1]);
    dszz_z = dszz_z + Dz3 * (wzz[r + 7] - wzz[r + 0]);
    const bool ok = col_ok && j < nz - 4;
    if (T && ok) {
      if (xk >= 0) {
        const size_t q = (size_t)j * 2 * T + xk;
        dsxx_x = cpml_apply(&d.psi[2][q], d.ax[i], d.bx[i], d.kx[i], dsxx_x);
        dsxz_x = cpml_apply(&d.psi[3][q], d.ax[i], d.bx[i], d.kx[i], dsxz_x);
      }
      const int zk = strip_x(j, nz, T, d.zlo, d.zhi);
      if (zk >= 0) {
        const size_t q = (size_t)zk * ldx + i;
        dsxz_z = cpml_apply(&d.psi[6][q], d.az[j], d.bz[j], d.kz[j], dsxz_z);
        dszz_z = cpml_apply(&d.psi[7][q], d.az[j], d.bz[j], d.kz[j], dszz_z);
      }
    }
    R fx = dsxx_x + dsxz_z;
    R fz = dszz_z + dsxz_x;
    for (int s = s0; s < s1; ++s) {
      if ((size_t)d.src_cell[s] != c) continue;
      const int kind = d.src_kind[s];
      const R val = step < d.snt ? d.src_w[(size_t)d.src_row[s] * d.snt + step] : R(0);
      if (kind == 0) fz = fz + val;
      else if (kind == 2) fx = fx + val;
    }
    const R bxv = two / (wrho[r] + wrho1[r]);
    const R bzv = two / (wrho[r] + wrho[r + 1]);
    if (ok) {
      vx[c] = ovx[r] + dt * (bxv * fx);
      vz[c] = ovz[r] + dt * (bzv * fz);
    }
  }
}

template <typename R, int L, int RZ>
__global__ __launch_bounds__(TX* NW) void k_stress(const ShotDesc<R>* __restrict__ shots,
                                                   const int* __restrict__ tile_shot,
                                                   int kstep) {
  __shared__ R sh_vx[NW][RZ][TX + 8];
  __shared__ R sh_vz[NW][RZ][TX + 8];
  const int lane = threadIdx.x, w = threadIdx.y;
  const int gtile = xcd_tile(blockIdx.x, gridDim.x);
  const ShotDesc<R>& d = shots[tile_shot[gtile]];
  const int tile = gtile - d.tile_base;
  const int ntx = d.ntx;
  const int tx = tile % ntx, tz = tile / ntx;
  const int i = 4 + tx * TX + lane;
  const int j0 = 4 + (tz * NW + w) * RZ;
  const int ldx = d.ldx, ldm = d.ldm;
  const size_t base = (size_t)j0 * ldx + i;
  const size_t mbase = (size_t)j0 * ldm + i;
  const R* __restrict__ vx = d.vx;
  const R* __restrict__ vz = d.vz;
  const R* __restrict__ mu = d.mu;
  const R* __restrict__ lam = d.lam;
  const R* __restrict__ tau = d.tau;
  R* __restrict__ sxx = d.sxx;
  R* __restrict__ szz = d.szz;
  R* __restrict__ sxz = d.sxz;
  R* __restrict__ M = d.M;
  const size_t plane = d.plane;

  R wvx[RZ + 7];  // vx rows j0-3 .. j0+RZ+3  (D+ along z, centre r+3)
  R wvz[RZ + 7];  // vz rows j0-4 .. j0+RZ+2  (D- along z, centre r+4)
  R wmu[RZ + 1], wmu1[RZ + 1];  // mu at columns i and i+1, rows j0 .. j0+RZ
#pragma unroll
  for (int k = 0; k < RZ + 7; ++k) {
    wvx[k] = vx[base + (ptrdiff_t)(k - 3) * ldx];
    wvz[k] = vz[base + (ptrdiff_t)(k - 4) * ldx];
  }
#pragma unroll
  for (int r = 0; r <= RZ; ++r) {
    wmu[r] = mu[mbase + (size_t)r * ldm];
    wmu1[r] = mu[mbase + (size_t)r * ldm + 1];
  }
  stage_rows<R, RZ, 3>(sh_vx[w], wvx, vx, base, ldx, lane);
  stage_rows<R, RZ, 4>(sh_vz[w], wvz, vz, base, ldx, lane);
  __syncthreads();

  const int nx = d.nx, nz = d.nz, T = d.T;
  const bool col_ok = i < nx - 4;
  const R dt = d.dt, two = R(2), quarter = R(0.25);
  const R Dx[4] = {d.Dx[0], d.Dx[1], d.Dx[2], d.Dx[3]};
  const R Dz0 = d.Dz[0], Dz1 = d.Dz[1], Dz2 = d.Dz[2], Dz3 = d.Dz[3];
  const int xk = T ? strip_x(i, nx, T, d.xlo, d.xhi) : -1;
  const int s0 = d.src_off[tile], s1 = d.src_off[tile + 1];
  const int r0 = d.rec_off[tile], r1 = d.rec_off[tile + 1];
  const long step = d.step + kstep;
  const int fs = d.fs;
  R a1[L > 0 ? L : 1], a2[L > 0 ? L : 1];
#pragma unroll
  for (int k = 0; k < L; ++k) {
    a1[k] = d.a1[k];
    a2[k] = d.a2[k];
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + r;
    const size_t c = base + (size_t)r * ldx;
    const size_t mc = mbase + (size_t)r * ldm;
    R exx = op_m(&sh_vx[w][r][lane + 4], 1, Dx);
    R exz = op_p(&sh_vz[w][r][lane + 4], 1, Dx);
    R ezz = Dz0 * (wvz[r + 4] - wvz[r + 3]);
    ezz = ezz + Dz1 * (wvz[r + 5] - wvz[r + 2]);
    ezz = ezz + Dz2 * (wvz[r + 6] - wvz[r + 1]);
    ezz = ezz + Dz3 * (wvz[r + 7] - wvz[r + 0]);
    R ezx = Dz0 * (wvx[r + 4] - wvx[r + 3]);
    ezx = ezx + Dz1 * (wvx[r + 5] - wvx[r + 2]);
    ezx = ezx + Dz2 * (wvx[r + 6] - wvx[r + 1]);
    ezx = ezx + Dz3 * (wvx[r + 7] - wvx[r + 0]);
    const bool ok = col_ok && j < nz - 4;
    if (!ok) continue;
    if (T) {
      if (xk >= 0) {
        const size_t q = (size_t)j * 2 * T + xk;
        exx = cpml_apply(&d.psi[0][q], d.ax[i], d.bx[i], d.kx[i], exx);
        exz = cpml_apply(&d.psi[1][q], d.ax[i], d.bx[i], d.kx[i], exz);
      }
      const int zk = strip_x(j, nz, T, d.zlo, d.zhi);
      if (zk >= 0) {
        const size_t q = (size_t)zk * ldx + i;
        ezx = cpml_apply(&d.psi[4][q], d.az[j], d.bz[j], d.kz[j], ezx);
        ezz = cpml_apply(&d.psi[5][q], d.az[j], d.bz[j], d.kz[j], ezz);
      }
    }
    const R gxz = ezx + exz;
    const R lr = lam[mc], mr = wmu[r];
    const R mxz = quarter * (((wmu[r] + wmu1[r]) + wmu[r + 1]) + wmu1[r + 1]);
    const R pr = lr + two * mr;
    const bool fs_row = fs && j == 4;
    R dsxx, dszz, dsxz;
    R mxx[L > 0 ? L : 1], mzz[L > 0 ? L : 1], mxzv[L > 0 ? L : 1];
    if constexpr (L > 0) {
      const R t = tau[mc];
      const R p1 = pr * t, p0 = p1 + pr;
      const R l1 = lr * t, l0 = l1 + lr;
      const R m1 = mxz * t, m0 = m1 + mxz;
#pragma unroll
      for (int k = 0; k < L; ++k) {
        mxx[k] = M[(size_t)k * plane + c];
        mzz[k] = M[(size_t)(L + k) * plane + c];
        mxzv[k] = M[(size_t)(2 * L + k) * plane + c];
      }
      R SMxx = mxx[0], SMzz = mzz[0], SMxz = mxzv[0];
#pragma unroll
      for (int k = 1; k < L; ++k) {
        SMxx = SMxx + mxx[k];
        SMzz = SMzz + mzz[k];
        SMxz = SMxz + mxzv[k];
      }
      if (fs_row) ezz = -(((l0 * exx) + l1 * SMxx) + p1 * SMzz) / p0;
      dsxx = p0 * exx + l0 * ezz + p1 * SMxx + l1 * SMzz;
      dszz = l0 * exx + p0 * ezz + l1 * SMxx + p1 * SMzz;
      dsxz = m0 * gxz + m1 * SMxz;
    } else {
      if (fs_row) ezz = -(lr * exx) / pr;
      dsxx = pr * exx + lr * ezz;
      dszz = lr * exx + pr * ezz;
      dsxz = mxz * gxz;
    }
    for (int s = s0; s < s1; ++s) {
      if ((size_t)d.src_cell[s] != c || d.src_kind[s] != 1) continue;
      const R val = step < d.snt ? d.src_w[(size_t)d.src_row[s] * d.snt + step] : R(0);
      dsxx = dsxx + val;
      dszz = dszz + val;
    }
    const R nsxx = sxx[c] + dt * dsxx;
    R nszz = szz[c] + dt * dszz;
    const R nsxz = sxz[c] + dt * dsxz;
    if (fs_row) nszz = R(0);
    sxx[c] = nsxx;
    szz[c] = nszz;
    sxz[c] = nsxz;
    if constexpr (L > 0) {
#pragma unroll
      for (int k = 0; k < L; ++k) {
        M[(size_t)k * plane + c] = mxx[k] + dt * (a1[k] * exx + a2[k] * mxx[k]);
        M[(size_t)(L + k) * plane + c] = mzz[k] + dt * (a1[k] * ezz + a2[k] * mzz[k]);
        M[(size_t)(2 * L + k) * plane + c] = mxzv[k] + dt * (a1[k] * gxz + a2[k] * mxzv[k]);
      }
    }
    if (fs) {  // stress images above the surface row 4
      if (j >= 5 && j <= 8) szz[(size_t)(8 - j) * ldx + i] = -nszz;
      if (j >= 4 && j <= 7) sxz[(size_t)(7 - j) * ldx + i] = -nsxz;
    }
    for (int q = r0; q < r1; ++q) {
      if ((size_t)d.rec_cell[q] != c || step >= d.rec_nt) continue;
      const int comp = d.rec_comp[q];
      const R val = comp == 0   ? sh_vx[w][r][lane + 4]
                    : comp == 1 ? sh_vz[w][r][lane + 4]
                    : comp == 2 ? nsxx
                    : comp == 3 ? nszz
                                : nsxz;
      d.traces[(size_t)step * d.nrec + d.rec_id[q]] = val;
    }
  }
}

// Advance the step index of every shot of a batch after a block of k steps.
template <typename R>
__global__ void k_advance(ShotDesc<R>* shots, int nshots, int k) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nshots) shots[s].step += k;
}

}  // namespace vi
