// vi_rates.hpp -- right-hand side of the step at the current state: the analogue of the
// reference's stress_eq!(ds, s, v, p, t) / velocity_eq!(dv, s, v, p, t)
// (src/sciml_solver/ViscoelasticProblem.jl:62-174) for the fused scheme of vi_kernels.hpp.
//
// k_rates writes, for every cell of [0, nz) x [0, nx) (0 outside the interior):
//   planes 0, 1     dvx, dvz: the velocity half step's rates from the stored stresses, force
//                   sources added before the buoyancy (values per source, not per step);
//   planes 2, 3, 4  dsxx, dszz, dsxz: the stress half step's rates from the stored velocities
//                   and memory variables, explosive sources added; on the free-surface row the
//                   ezz that makes dszz = 0 is used and dszz is 0 (with formula #3, row 5
//                   recomputes that value from row 4's operands, which k_rates leaves as is);
//   planes 5 ..     dM (3 L planes: xx, zz, xz per mechanism): a1_n e + a2_n M_n.
// In CPML strips a derivative uses psi' = b psi + a d without storing it, as the step does, so
// the step's update is x + dt * rate operation for operation (checked against the oracle).
// k_rates_images then gives the stress-imaging rows the negated rates of their mirror rows.
// Not a hot path: one thread per cell, operands read from the padded planes directly.
#pragma once
#include "vi_kernels.hpp"

namespace vi {

// out: rate planes of the state pitch ldx, `plane` elements apart; sources: cell (six, siz),
// kind and value of each source in the caller's order (sval null: no source term)
template <typename R, int L>
__global__ __launch_bounds__(256) void k_rates(const ShotDesc<R> d, const int* __restrict__ six,
                                               const int* __restrict__ siz,
                                               const int* __restrict__ skind,
                                               const R* __restrict__ sval, int nsrc,
                                               R* __restrict__ out) {
  constexpr unsigned E = sizeof(R);
  constexpr int LM = L > 0 ? L : 1;
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx || j >= d.nz) return;
  const int ldx = d.ldx, nx = d.nx, nz = d.nz, T = d.T;
  const size_t P = d.plane;
  const int c = j * ldx + i;
  const unsigned cb = (unsigned)c * E;
  const bool in = i >= 4 && i < nx - 4 && j >= 4 && j < nz - 4;
  if (!in) {
    for (int q = 0; q < 5 + 3 * L; ++q) gst(out + q * P, cb, R(0));
    return;
  }
  const R two = R(2), quarter = R(0.25);
  const R Dx[4] = {d.Dx[0], d.Dx[1], d.Dx[2], d.Dx[3]};
  const R Dz[4] = {d.Dz[0], d.Dz[1], d.Dz[2], d.Dz[3]};
  const bool afda = d.fs >= 2;
  const int xk = T ? strip_x(i, nx, T, d.xlo, d.xhi) : -1;
  const int zk = T ? strip_x(j, nz, T, d.zlo, d.zhi) : -1;
  const unsigned qx = (unsigned)(j * 2 * T + xk) * E, qz = (unsigned)(zk * ldx + i) * E;
  // ---- velocity rates
  R dsxx_x = op_p(d.sxx + c, 1, Dx);
  R dsxz_z = afda && j < 8 ? afda_dz(d.sxz + 4 * ldx + i, ldx, d.afda, j - 4)
                           : op_m(d.sxz + c, ldx, Dz);
  R dsxz_x = op_m(d.sxz + c, 1, Dx);
  R dszz_z = afda && j < 7 ? afda_dz(d.szz + 4 * ldx + i, ldx, d.afda, j)
                           : op_p(d.szz + c, ldx, Dz);
  if (xk >= 0) {
    const R a = d.ax[i], b = d.bx[i], k = d.kx[i];
    dsxx_x = cpml_rate(d.psi[2], qx, a, b, k, dsxx_x);
    dsxz_x = cpml_rate(d.psi[3], qx, a, b, k, dsxz_x);
  }
  if (zk >= 0) {
    const R a = d.az[j], b = d.bz[j], k = d.kz[j];
    dsxz_z = cpml_rate(d.psi[6], qz, a, b, k, dsxz_z);
    dszz_z = cpml_rate(d.psi[7], qz, a, b, k, dszz_z);
  }
  R fx = dsxx_x + dsxz_z;
  R fz = dszz_z + dsxz_x;
  for (int q = 0; q < nsrc; ++q) {
    if (six[q] != i || siz[q] != j) continue;
    const R w = sval ? sval[q] : R(0);
    if (skind[q] == 0) fz = fz + w;
    else if (skind[q] == 2) fx = fx + w;
  }
  const int ldm = d.ldm;
  const int m = j * ldm + i;
  gst(out, cb, (two / (d.rho[m] + d.rho[m + 1])) * fx);
  gst(out + P, cb, (two / (d.rho[m] + d.rho[m + ldm])) * fz);
  // ---- stress rates
  R exx = op_m(d.vx + c, 1, Dx);
  R exz = op_p(d.vz + c, 1, Dx);
  R ezz = d.fs == 3 && j == 5
              ? herm_dz(d.vz + 4 * ldx + i, ldx, d.afda,
                        surface_ezz<R, L>(d, i, op_m(d.vx + 4 * ldx + i, 1, Dx)))
          : afda && j >= 5 && j < 8 ? afda_dz(d.vz + 4 * ldx + i, ldx, d.afda, j + 2)
                                    : op_m(d.vz + c, ldx, Dz);
  R ezx = afda && j < 7 ? afda_dz(d.vx + 4 * ldx + i, ldx, d.afda, j + 6)
                        : op_p(d.vx + c, ldx, Dz);
  if (xk >= 0) {
    const R a = d.ax[i], b = d.bx[i], k = d.kx[i];
    exx = cpml_rate(d.psi[0], qx, a, b, k, exx);
    exz = cpml_rate(d.psi[1], qx, a, b, k, exz);
  }
  if (zk >= 0) {
    const R a = d.az[j], b = d.bz[j], k = d.kz[j];
    ezx = cpml_rate(d.psi[4], qz, a, b, k, ezx);
    ezz = cpml_rate(d.psi[5], qz, a, b, k, ezz);
  }
  const R gxz = ezx + exz;
  const R lr = d.lam[m];
  const R mxz = quarter * (((d.mu[m] + d.mu[m + 1]) + d.mu[m + ldm]) + d.mu[m + ldm + 1]);
  const R px = d.vti ? d.c11[m] : lr + two * d.mu[m];
  const R pz = d.vti ? d.c33[m] : px;
  const bool fs_row = d.fs && j == 4;
  R rxx, rzz, rxz;
  if constexpr (L > 0) {
    const R t = d.tau[m];
    const R px1 = px * t, px0 = px1 + px;
    const R pz1 = pz * t, pz0 = pz1 + pz;
    const R l1 = lr * t, l0 = l1 + lr;
    const R m1 = mxz * t, m0 = m1 + mxz;
    R Mv[3][LM];
#pragma unroll
    for (int k = 0; k < L; ++k)
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) Mv[comp][k] = gld(d.M + (comp * L + k) * P, cb);
    R SMxx = Mv[0][0], SMzz = Mv[1][0], SMxz = Mv[2][0];
#pragma unroll
    for (int k = 1; k < L; ++k) {
      SMxx = SMxx + Mv[0][k];
      SMzz = SMzz + Mv[1][k];
      SMxz = SMxz + Mv[2][k];
    }
    if (fs_row) ezz = -(((l0 * exx) + l1 * SMxx) + pz1 * SMzz) / pz0;
    rxx = px0 * exx + l0 * ezz + px1 * SMxx + l1 * SMzz;
    rzz = l0 * exx + pz0 * ezz + l1 * SMxx + pz1 * SMzz;
    rxz = m0 * gxz + m1 * SMxz;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const R a1 = d.a1[k], a2 = d.a2[k];
      gst(out + (5 + k) * P, cb, a1 * exx + a2 * Mv[0][k]);
      gst(out + (5 + L + k) * P, cb, a1 * ezz + a2 * Mv[1][k]);
      gst(out + (5 + 2 * L + k) * P, cb, a1 * gxz + a2 * Mv[2][k]);
    }
  } else {
    if (fs_row) ezz = -(lr * exx) / pz;
    rxx = px * exx + lr * ezz;
    rzz = lr * exx + pz * ezz;
    rxz = mxz * gxz;
  }
  for (int q = 0; q < nsrc; ++q) {  // explosive sources, in the caller's order
    if (six[q] != i || siz[q] != j || skind[q] != 1) continue;
    const R w = sval ? sval[q] : R(0);
    rxx = rxx + w;
    rzz = rzz + w;
  }
  gst(out + 2 * P, cb, rxx);
  gst(out + 3 * P, cb, fs_row ? R(0) : rzz);
  gst(out + 4 * P, cb, rxz);
}

// stress imaging: rows 4 - m (m = 1..4) carry -dszz of row 4 + m and -dsxz of row 3 + m
template <typename R>
__global__ void k_rates_images(int nx, int ldx, size_t plane, R* __restrict__ out) {
  const int i = 4 + blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 1 + threadIdx.y;
  if (i >= nx - 4) return;
  R* dszz = out + 3 * plane;
  R* dsxz = out + 4 * plane;
  dszz[(4 - m) * ldx + i] = -dszz[(4 + m) * ldx + i];
  dsxz[(4 - m) * ldx + i] = -dsxz[(4 + m - 1) * ldx + i];
}

// Whether every interior value of the planes is finite: a thread that finds a NaN or an
// infinity stores 1 into *flag (cleared before; every writer writes the same value).
template <typename R>
__global__ __launch_bounds__(256) void k_nonfinite(const R* __restrict__ a,
                                                   const R* __restrict__ b, int nx, int nz,
                                                   int ldx, int* __restrict__ flag) {
  const int i = 4 + blockIdx.x * 64 + threadIdx.x;
  const int j = 4 + blockIdx.y * 4 + threadIdx.y;
  if (i >= nx - 4 || j >= nz - 4) return;
  const size_t c = (size_t)j * ldx + i;
  if (!isfinite(a[c]) || !isfinite(b[c])) *flag = 1;
}

}  // namespace vi
