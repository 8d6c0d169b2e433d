// vi_kernels.hpp -- gfx950 kernels of the fused velocity-stress step.
//
// One step = k_velocity (vx, vz from the stress divergence) then k_stress (sxx, szz, sxz
// and the SLS memory variables from the velocity gradients). Each launch covers the
// interior [4, nx-4) x [4, nz-4) with tiles of TX x (NW*RZ) cells:
//   - a workgroup is NW wavefronts of 64 lanes; lane = x offset, wave = band of RZ rows;
//   - fields differentiated along x are staged row by row in LDS with a 4-column halo;
//   - fields differentiated along z are held as a register window of RZ+7 rows per lane
//     (z register blocking), so every global load is a coalesced 64-lane row segment;
//   - RZ (rows per wave) is a template parameter chosen per grid by the autotuner: small
//     grids need small RZ for enough waves in flight, large grids amortise the z halo;
//   - the workgroup -> tile map is XCD-aware: consecutive tiles of a row band go to the
//     same XCD so vertical halos are re-read from that XCD's L2.
// Arrays are padded (pitch ldx >= nx + 72, plane rows >= nz + NW*8 + 16) so that halo and
// window loads of edge tiles stay inside the allocation without bounds checks; only
// stores are guarded. The padding is zero and is never written.
//
// Arithmetic follows the oracle's operation order exactly (no contraction: the library is
// built with -ffp-contract=off), so the FP32 and FP64 results match the CPU oracle bit for
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

template <typename R>
struct StepParams {
  int nx, nz, ldx;
  int ntx, ntiles;
  size_t plane;  // elements per padded field plane
  R dt;
  R Dx[4], Dz[4];
  R *vx, *vz, *sxx, *szz, *sxz;
  R* M;  // [3][L] planes
  const R *rho, *lam, *mu, *tau;
  R a1[MAXL], a2[MAXL];
  int T, xlo, xhi, zlo, zhi;
  const R *ax, *bx, *kx, *az, *bz, *kz;
  R* psi[8];  // 0..3: [nz][2T] (vx_x, vz_x, sxx_x, sxz_x); 4..7: [2T][ldx] (vx_z, vz_z, sxz_z, szz_z)
  int fs;
  // sources, grouped by tile (CSR), original order kept within a tile
  const int* src_off;
  const int* src_cell;
  const int* src_kind;
  const R* src_w;  // [n][snt], row = original source index
  const int* src_row;
  int snt;
  // receivers, grouped by tile
  const int* rec_off;
  const int* rec_cell;
  const int* rec_comp;
  const int* rec_id;
  R* traces;  // [rec_nt][nrec]
  int nrec, rec_nt;
  long step;
};

// Workgroup id -> tile id such that each of the 8 XCDs (which receive workgroups round
// robin) owns one contiguous range of tiles.
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
__global__ __launch_bounds__(TX* NW) void k_velocity(const StepParams<R> p) {
  __shared__ R sh_xx[NW][RZ][TX + 8];
  __shared__ R sh_xz[NW][RZ][TX + 8];
  const int lane = threadIdx.x, w = threadIdx.y;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % p.ntx, tz = tile / p.ntx;
  const int i = 4 + tx * TX + lane;
  const int j0 = 4 + (tz * NW + w) * RZ;
  const int ldx = p.ldx;
  const size_t base = (size_t)j0 * ldx + i;
  const R* __restrict__ sxx = p.sxx;
  const R* __restrict__ sxz = p.sxz;
  const R* __restrict__ szz = p.szz;
  const R* __restrict__ rho = p.rho;
  R* __restrict__ vx = p.vx;
  R* __restrict__ vz = p.vz;

  R wxx[RZ];      // sxx rows j0 .. j0+RZ-1
  R wxz[RZ + 7];  // sxz rows j0-4 .. j0+RZ+2   (D- along z, centre at r+4)
  R wzz[RZ + 7];  // szz rows j0-3 .. j0+RZ+3   (D+ along z, centre at r+3)
  R wrho[RZ + 1], wrho1[RZ];  // rho rows j0 .. j0+RZ at column i; rows j0.. at column i+1
  R ovx[RZ], ovz[RZ];
#pragma unroll
  for (int k = 0; k < RZ + 7; ++k) {
    wxz[k] = sxz[base + (ptrdiff_t)(k - 4) * ldx];
    wzz[k] = szz[base + (ptrdiff_t)(k - 3) * ldx];
  }
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    wxx[r] = sxx[base + (size_t)r * ldx];
    wrho1[r] = rho[base + (size_t)r * ldx + 1];
    ovx[r] = vx[base + (size_t)r * ldx];
    ovz[r] = vz[base + (size_t)r * ldx];
  }
#pragma unroll
  for (int r = 0; r <= RZ; ++r) wrho[r] = rho[base + (size_t)r * ldx];

  stage_rows<R, RZ, 0>(sh_xx[w], wxx, sxx, base, ldx, lane);
  stage_rows<R, RZ, 4>(sh_xz[w], wxz, sxz, base, ldx, lane);
  __syncthreads();

  const bool col_ok = i < p.nx - 4;
  const R dt = p.dt, two = R(2);
  const R Dz0 = p.Dz[0], Dz1 = p.Dz[1], Dz2 = p.Dz[2], Dz3 = p.Dz[3];
  const int xk = p.T ? strip_x(i, p.nx, p.T, p.xlo, p.xhi) : -1;
  const int s0 = p.src_off[tile], s1 = p.src_off[tile + 1];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + r;
    const size_t c = base + (size_t)r * ldx;
    R dsxx_x = op_p(&sh_xx[w][r][lane + 4], 1, p.Dx);
    R dsxz_x = op_m(&sh_xz[w][r][lane + 4], 1, p.Dx);
    // D- of sxz along z (centre r+4) and D+ of szz along z (centre r+3)
    R dsxz_z = Dz0 * (wxz[r + 4] - wxz[r + 3]);
    dsxz_z = dsxz_z + Dz1 * (wxz[r + 5] - wxz[r + 2]);
    dsxz_z = dsxz_z + Dz2 * (wxz[r + 6] - wxz[r + 1]);
    dsxz_z = dsxz_z + Dz3 * (wxz[r + 7] - wxz[r + 0]);
    R dszz_z = Dz0 * (wzz[r + 4] - wzz[r + 3]);
    dszz_z = dszz_z + Dz1 * (wzz[r + 5] - wzz[r + 2]);
    dszz_z = dszz_z + Dz2 * (wzz[r + 6] - wzz[r + 1]);
    dszz_z = dszz_z + Dz3 * (wzz[r + 7] - wzz[r + 0]);
    const bool ok = col_ok && j < p.nz - 4;
    if (p.T && ok) {
      if (xk >= 0) {
        const size_t q = (size_t)j * 2 * p.T + xk;
        dsxx_x = cpml_apply(&p.psi[2][q], p.ax[i], p.bx[i], p.kx[i], dsxx_x);
        dsxz_x = cpml_apply(&p.psi[3][q], p.ax[i], p.bx[i], p.kx[i], dsxz_x);
      }
      const int zk = strip_x(j, p.nz, p.T, p.zlo, p.zhi);
      if (zk >= 0) {
        const size_t q = (size_t)zk * ldx + i;
        dsxz_z = cpml_apply(&p.psi[6][q], p.az[j], p.bz[j], p.kz[j], dsxz_z);
        dszz_z = cpml_apply(&p.psi[7][q], p.az[j], p.bz[j], p.kz[j], dszz_z);
      }
    }
    R fx = dsxx_x + dsxz_z;
    R fz = dszz_z + dsxz_x;
    for (int s = s0; s < s1; ++s) {
      if ((size_t)p.src_cell[s] != c) continue;
      const int kind = p.src_kind[s];
      const R val = p.step < p.snt ? p.src_w[(size_t)p.src_row[s] * p.snt + p.step] : R(0);
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
__global__ __launch_bounds__(TX* NW) void k_stress(const StepParams<R> p) {
  __shared__ R sh_vx[NW][RZ][TX + 8];
  __shared__ R sh_vz[NW][RZ][TX + 8];
  const int lane = threadIdx.x, w = threadIdx.y;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % p.ntx, tz = tile / p.ntx;
  const int i = 4 + tx * TX + lane;
  const int j0 = 4 + (tz * NW + w) * RZ;
  const int ldx = p.ldx;
  const size_t base = (size_t)j0 * ldx + i;
  const R* __restrict__ vx = p.vx;
  const R* __restrict__ vz = p.vz;
  const R* __restrict__ mu = p.mu;
  const R* __restrict__ lam = p.lam;
  const R* __restrict__ tau = p.tau;
  R* __restrict__ sxx = p.sxx;
  R* __restrict__ szz = p.szz;
  R* __restrict__ sxz = p.sxz;
  R* __restrict__ M = p.M;
  const size_t plane = p.plane;

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
    wmu[r] = mu[base + (size_t)r * ldx];
    wmu1[r] = mu[base + (size_t)r * ldx + 1];
  }
  stage_rows<R, RZ, 3>(sh_vx[w], wvx, vx, base, ldx, lane);
  stage_rows<R, RZ, 4>(sh_vz[w], wvz, vz, base, ldx, lane);
  __syncthreads();

  const bool col_ok = i < p.nx - 4;
  const R dt = p.dt, two = R(2), quarter = R(0.25);
  const R Dz0 = p.Dz[0], Dz1 = p.Dz[1], Dz2 = p.Dz[2], Dz3 = p.Dz[3];
  const int xk = p.T ? strip_x(i, p.nx, p.T, p.xlo, p.xhi) : -1;
  const int s0 = p.src_off[tile], s1 = p.src_off[tile + 1];
  const int r0 = p.rec_off[tile], r1 = p.rec_off[tile + 1];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + r;
    const size_t c = base + (size_t)r * ldx;
    R exx = op_m(&sh_vx[w][r][lane + 4], 1, p.Dx);
    R exz = op_p(&sh_vz[w][r][lane + 4], 1, p.Dx);
    R ezz = Dz0 * (wvz[r + 4] - wvz[r + 3]);
    ezz = ezz + Dz1 * (wvz[r + 5] - wvz[r + 2]);
    ezz = ezz + Dz2 * (wvz[r + 6] - wvz[r + 1]);
    ezz = ezz + Dz3 * (wvz[r + 7] - wvz[r + 0]);
    R ezx = Dz0 * (wvx[r + 4] - wvx[r + 3]);
    ezx = ezx + Dz1 * (wvx[r + 5] - wvx[r + 2]);
    ezx = ezx + Dz2 * (wvx[r + 6] - wvx[r + 1]);
    ezx = ezx + Dz3 * (wvx[r + 7] - wvx[r + 0]);
    const bool ok = col_ok && j < p.nz - 4;
    if (!ok) continue;
    if (p.T) {
      if (xk >= 0) {
        const size_t q = (size_t)j * 2 * p.T + xk;
        exx = cpml_apply(&p.psi[0][q], p.ax[i], p.bx[i], p.kx[i], exx);
        exz = cpml_apply(&p.psi[1][q], p.ax[i], p.bx[i], p.kx[i], exz);
      }
      const int zk = strip_x(j, p.nz, p.T, p.zlo, p.zhi);
      if (zk >= 0) {
        const size_t q = (size_t)zk * ldx + i;
        ezx = cpml_apply(&p.psi[4][q], p.az[j], p.bz[j], p.kz[j], ezx);
        ezz = cpml_apply(&p.psi[5][q], p.az[j], p.bz[j], p.kz[j], ezz);
      }
    }
    const R gxz = ezx + exz;
    const R lr = lam[c], mr = wmu[r];
    const R mxz = quarter * (((wmu[r] + wmu1[r]) + wmu[r + 1]) + wmu1[r + 1]);
    const R pr = lr + two * mr;
    const bool fs_row = p.fs && j == 4;
    R dsxx, dszz, dsxz;
    R mxx[L > 0 ? L : 1], mzz[L > 0 ? L : 1], mxzv[L > 0 ? L : 1];
    if constexpr (L > 0) {
      const R t = tau[c];
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
      if ((size_t)p.src_cell[s] != c || p.src_kind[s] != 1) continue;
      const R val = p.step < p.snt ? p.src_w[(size_t)p.src_row[s] * p.snt + p.step] : R(0);
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
        const R a1 = p.a1[k], a2 = p.a2[k];
        M[(size_t)k * plane + c] = mxx[k] + dt * (a1 * exx + a2 * mxx[k]);
        M[(size_t)(L + k) * plane + c] = mzz[k] + dt * (a1 * ezz + a2 * mzz[k]);
        M[(size_t)(2 * L + k) * plane + c] = mxzv[k] + dt * (a1 * gxz + a2 * mxzv[k]);
      }
    }
    if (p.fs) {  // stress images above the surface row 4
      if (j >= 5 && j <= 8) szz[(size_t)(8 - j) * ldx + i] = -nszz;
      if (j >= 4 && j <= 7) sxz[(size_t)(7 - j) * ldx + i] = -nsxz;
    }
    for (int q = r0; q < r1; ++q) {
      if ((size_t)p.rec_cell[q] != c || p.step >= p.rec_nt) continue;
      const int comp = p.rec_comp[q];
      const R val = comp == 0   ? sh_vx[w][r][lane + 4]
                    : comp == 1 ? sh_vz[w][r][lane + 4]
                    : comp == 2 ? nsxx
                    : comp == 3 ? nszz
                                : nsxz;
      p.traces[(size_t)p.step * p.nrec + p.rec_id[q]] = val;
    }
  }
}

}  // namespace vi
