// vi_adjoint.hpp -- discrete adjoint of the fused velocity-stress step.
//
// The forward step n is u^{n+1} = S(V(u^n)) with V the velocity half step and S the stress
// half step (k_velocity, k_stress). For a fixed model both are affine in the state
// u = (v, s, M, psi), so the gradient of J = 1/2 sum_n |R u^{n+1} - d_obs^n|^2 follows from
// the transposed half steps applied backwards in time:
//   for n = nt-1 .. 0:  u_hat += R^T r^n;  u_hat = S^T u_hat;  u_hat = V^T u_hat
// with the model gradient accumulated as u_hat_post . d(half step)/d(model) . u_pre.
// Each transposed half step is a pointwise kernel (adjoint of the constitutive / buoyancy /
// CPML / memory-variable arithmetic of one cell, which also writes the adjoint of every raw
// derivative) followed by a gather kernel that applies the transposed 8th-order operators:
// (D+)^T = -D- and (D-)^T = -D+ on the padded interior, so s_hat += -D-(d_hat) etc.
// The forward quantities each adjoint kernel needs (the CPML-stretched strain rates, the stress
// divergence and the memory-variable sums of that step) are stored by the recording forward
// kernels, so the transposed pointwise kernels read them per cell without halos.
// The forward scheme is the one of src/visco-kernels.jl:3-85 with the CPML of
// src/sciml_solver/CPML.jl:291-328 and the stress imaging of CPML.jl:372-380 (fs = 1) or the
// H-AFDA rows of src/free-surface-kernels.jl (fs = 2). With H-AFDA the raw z-derivative G of
// the rows next to the surface is a one-sided stencil G_a instead of the centred G_c, so
// G^T e = G_c^T e + (G_a - G_c)^T e_A over those rows (e_A): the gathers add the second term,
// from a table of G_a - G_c for the target rows 4..11, to the centred transposed stencils.
#pragma once
#include "vi_kernels.hpp"

namespace vi {

// Everything the adjoint kernels of one shot need.
template <typename R>
struct AdjDesc {
  int nx, nz, ldx, ldm;
  R dt, Dx[4], Dz[4];
  R a1[MAXL], a2[MAXL];
  int L, T, xlo, xhi, zlo, zhi, fs;
  // fs = 2: H-AFDA weights [13][8] (rows A0..3 B0..2 C1..3 D0..2 over grid rows 4..11),
  // followed by [13][8] of G_a - G_c (the same source rows, target rows 4..11)
  const R* afda;
  size_t plane;
  const R *rho, *lam, *mu, *tau;  // model (pitch ldm); VTI: lam = C13, mu = C55
  const R *c11, *c33;             // VTI only
  int vti;
  const R *ax, *bx, *kx, *az, *bz, *kz;
  // forward quantities of the current step (pitch ldx, see RecSlot): the velocity half's
  // stress divergence fx, fz (stretched, sources included, before the buoyancy), the stress
  // half's stretched strain rates exx, ezz, gxz and the memory-variable sums sum_k M^n_k of
  // the xx, zz and xz components (three planes)
  const R *ffx, *ffz, *fexx, *fezz, *fgxz, *fM;
  // adjoint state (pitch ldx) and per-step work arrays
  R *avx, *avz, *asxx, *aszz, *asxz, *aM;
  R* apsi[8];
  R* ehat[4];  // adjoint of raw D-x vx, D+z vx, D-z vz, D+x vz (stress half)
  R* dhat[4];  // adjoint of raw D+x sxx, D-z sxz, D-x sxz, D+z szz (velocity half)
  // gradient accumulators (pitch ldx): lam, mu (through C11 and C33), mu_xz, tau, bx, bz;
  // VTI: g_lam = C13, g_mu = C11, g_c33 = C33, g_muxz = C55 at the sxz node
  R *g_lam, *g_mu, *g_muxz, *g_tau, *g_bx, *g_bz, *g_c33;
  // residual injection into the velocity adjoint, done by k_adj_stress (which does not touch
  // v_hat): the first ntgt_v targets of the CSR target list (components vx, vz), and the
  // residual row of this step
  const int *tcell, *tcomp, *tptr, *trec;
  int ntgt_v;
  const R* resid_row;
};

template <typename R>
__device__ __forceinline__ R gp(const R* g, unsigned cb, int stb, const R* D) {
  R v = D[0] * (gld(g, cb + stb) - gld(g, cb));
  v = v + D[1] * (gld(g, cb + 2 * stb) - gld(g, cb - stb));
  v = v + D[2] * (gld(g, cb + 3 * stb) - gld(g, cb - 2 * stb));
  v = v + D[3] * (gld(g, cb + 4 * stb) - gld(g, cb - 3 * stb));
  return v;
}
template <typename R>
__device__ __forceinline__ R gm(const R* g, unsigned cb, int stb, const R* D) {
  R v = D[0] * (gld(g, cb) - gld(g, cb - stb));
  v = v + D[1] * (gld(g, cb + stb) - gld(g, cb - 2 * stb));
  v = v + D[2] * (gld(g, cb + 2 * stb) - gld(g, cb - 3 * stb));
  v = v + D[3] * (gld(g, cb + 3 * stb) - gld(g, cb - 4 * stb));
  return v;
}

// adjoint of (p' = b p + a d; dt = d/k + p'): given the adjoint of the stretched value (h)
// and of p' (ph), returns the adjoint of d and writes the adjoint of p into *ph
template <typename R>
__device__ __forceinline__ R cpml_adj(R h, R* ph, R a, R b, R k) {
  const R tot = *ph + h;
  *ph = b * tot;
  return h / k + a * tot;
}

// ---- adjoint of the stress half step (pointwise part of one interior cell) ------------
// exx = D-x vx, ezx = D+z vx, ezz = D-z vz, exz = D+x vz: the raw forward rates of the cell
// Everything one cell of the transposed stress half reads from memory: forward CPML state and
// model, stored memory sums, and every value the cell reads back (adjoint state, memory-
// variable adjoints, gradient accumulators, CPML adjoints, image rows). The kernels request it
// for all rows of a wavefront before the tile barrier, so that the whole tile makes one memory
// round trip: the planes are distinct, but the compiler cannot prove it and would otherwise
// wait for each read-modify-write in turn. Only for interior cells.
template <typename R, int L>
struct AdjStressIn {
  static constexpr int LM = L > 0 ? L : 1;
  R axv = 0, bxv = 0, kxv = 1, azv = 0, bzv = 0, kzv = 1;
  R p0h = 0, p1h = 0, p4h = 0, p5h = 0;
  R exx, ezz, gxz;  // the forward step's stretched strain rates
  R lr, mr, mu1, mu2, mu3, c11v, c33v, t, SMxx = 0, SMzz = 0, SMxz = 0;
  R sxxh, szz_c, sxz_c, szz_im, sxz_im;
  R mh[3 * LM];
  R acc_lam, acc_mu, acc_c33, acc_muxz, acc_tau;
  __device__ __forceinline__ void load(const AdjDesc<R>& d, int i, int j) {
    constexpr unsigned E = sizeof(R);
    const int ldx = d.ldx;
    const unsigned cb = (unsigned)(j * ldx + i) * E;
    const unsigned mb = (unsigned)(j * d.ldm + i) * E;
    const int xk = d.T ? strip_x(i, d.nx, d.T, d.xlo, d.xhi) : -1;
    const int zk = d.T ? strip_x(j, d.nz, d.T, d.zlo, d.zhi) : -1;
    const unsigned qx = (unsigned)(j * 2 * d.T + xk) * E, qz = (unsigned)(zk * ldx + i) * E;
    if (xk >= 0) {
      axv = gld(d.ax, i * E); bxv = gld(d.bx, i * E); kxv = gld(d.kx, i * E);
      p0h = gld((const R*)d.apsi[0], qx); p1h = gld((const R*)d.apsi[1], qx);
    }
    if (zk >= 0) {
      azv = d.az[j]; bzv = d.bz[j]; kzv = d.kz[j];
      p4h = gld((const R*)d.apsi[4], qz); p5h = gld((const R*)d.apsi[5], qz);
    }
    exx = gld_nt(d.fexx, cb);
    ezz = gld_nt(d.fezz, cb);
    gxz = gld_nt(d.fgxz, cb);
    lr = gld(d.lam, mb);
    mr = gld(d.mu, mb);
    mu1 = gld(d.mu, mb + E);
    mu2 = gld(d.mu, mb + d.ldm * E);
    mu3 = gld(d.mu, mb + d.ldm * E + E);
    c11v = d.vti ? gld(d.c11, mb) : R(0);
    c33v = d.vti ? gld(d.c33, mb) : R(0);
    t = L > 0 ? gld(d.tau, mb) : R(0);
    if (L > 0) {
      SMxx = gld_nt(d.fM, cb);
      SMzz = gld_nt(d.fM + d.plane, cb);
      SMxz = gld_nt(d.fM + 2 * d.plane, cb);
    }
    sxxh = gld((const R*)d.asxx, cb);
    szz_c = gld((const R*)d.aszz, cb);
    sxz_c = gld((const R*)d.asxz, cb);
    const bool imz = d.fs == 1 && j >= 5 && j <= 8, imx = d.fs == 1 && j >= 4 && j <= 7;
    szz_im = imz ? gld((const R*)d.aszz, (unsigned)((8 - j) * ldx + i) * E) : R(0);
    sxz_im = imx ? gld((const R*)d.asxz, (unsigned)((7 - j) * ldx + i) * E) : R(0);
    for (int k = 0; k < 3 * L; ++k) mh[k] = gld((const R*)(d.aM + (size_t)k * d.plane), cb);
    acc_lam = gld((const R*)d.g_lam, cb);
    acc_mu = gld((const R*)d.g_mu, cb);
    acc_c33 = d.vti ? gld((const R*)d.g_c33, cb) : R(0);
    acc_muxz = gld((const R*)d.g_muxz, cb);
    acc_tau = L > 0 ? gld((const R*)d.g_tau, cb) : R(0);
  }
};

// esh: with formula #3 (fs = 3), the adjoint of the surface value es that row 5's d vz/dz
// reads (alpha times row 5's raw-rate adjoint), added on the surface row (j = 4), whose
// free-surface rate ezz* is that same value; zero otherwise. Returns the adjoint of the raw
// d vz/dz of the cell (what formula #3's row 5 hands to the surface row).
// HERM: the variant for formula #3 (fs = 3), the only one that reads esh.
template <typename R, int L, bool HERM = false>
__device__ __forceinline__ R adj_stress_cell(const AdjDesc<R>& d, const AdjStressIn<R, L>& in,
                                             int i, int j, R esh = R(0)) {
  constexpr unsigned E = sizeof(R);
  constexpr int LM = L > 0 ? L : 1;
  const int ldx = d.ldx;
  const unsigned cb = (unsigned)(j * ldx + i) * E;
  const R dt = d.dt, two = R(2), quarter = R(0.25);
  const int xk = d.T ? strip_x(i, d.nx, d.T, d.xlo, d.xhi) : -1;
  const int zk = d.T ? strip_x(j, d.nz, d.T, d.zlo, d.zhi) : -1;
  const unsigned qx = (unsigned)(j * 2 * d.T + xk) * E, qz = (unsigned)(zk * ldx + i) * E;
  const bool fs_row = d.fs && j == 4;
  const R axv = in.axv, bxv = in.bxv, kxv = in.kxv, azv = in.azv, bzv = in.bzv, kzv = in.kzv;
  R p0h = in.p0h, p1h = in.p1h, p4h = in.p4h, p5h = in.p5h;
  const R lr = in.lr, mr = in.mr, t = in.t, SMxx = in.SMxx, SMzz = in.SMzz, SMxz = in.SMxz;
  const R sxxh = in.sxxh;
  R szz_c = in.szz_c, sxz_c = in.sxz_c;
  const bool imz = d.fs == 1 && j >= 5 && j <= 8, imx = d.fs == 1 && j >= 4 && j <= 7;
  const unsigned ibz = (unsigned)((8 - j) * ldx + i) * E, ibx = (unsigned)((7 - j) * ldx + i) * E;
  R* Ma[3 * LM];
  for (int k = 0; k < 3 * L; ++k) Ma[k] = d.aM + (size_t)k * d.plane;
  // ---- forward strain rates of this step (CPML-stretched), stored by the recording kernel
  const R exx = in.exx, ezz = in.ezz, gxz = in.gxz;
  const R mxz = quarter * (((mr + in.mu1) + in.mu2) + in.mu3);
  const R px = d.vti ? in.c11v : lr + two * mr;
  const R pz = d.vti ? in.c33v : px;
  const R px1 = px * t, px0 = px1 + px, pz1 = pz * t, pz0 = pz1 + pz;
  const R l1 = lr * t, l0 = l1 + lr, m1 = mxz * t, m0 = m1 + mxz;
  const R N = l0 * exx + l1 * SMxx + pz1 * SMzz;  // free-surface row: ezz* = -N/C33_0
  const R ezs = fs_row ? -N / pz0 : ezz;
  // stress images: their adjoint goes to the interior rows they copy, then they are cleared
  if (imz) {
    szz_c = szz_c - in.szz_im;
    gst(d.aszz, cb, szz_c);
    gst(d.aszz, ibz, R(0));
  }
  if (imx) {
    sxz_c = sxz_c - in.sxz_im;
    gst(d.asxz, cb, sxz_c);
    gst(d.asxz, ibx, R(0));
  }
  const R szzh = fs_row ? R(0) : szz_c;
  const R sxzh = sxz_c;
  if (fs_row) gst(d.aszz, cb, R(0));
  // adjoints of the (stretched) rates and of the memory sums
  R exh = dt * (px0 * sxxh + l0 * szzh);
  R ezh = dt * (l0 * sxxh + pz0 * szzh);
  R gh = dt * (m0 * sxzh);
  const R SMxxh = dt * (px1 * sxxh + l1 * szzh);
  const R SMzzh = dt * (l1 * sxxh + pz1 * szzh);
  const R SMxzh = dt * (m1 * sxzh);
  R Mzzh_sum = 0;  // sum_n a1_n * M_hat'_zz,n (free-surface gradient term)
  R fsMxx = 0, fsMzz = 0;  // extra adjoint of each M_n through ezz* on the free-surface row
  R ezh_from_M = 0;
  R nm[3 * LM];  // the new memory-variable adjoints
  for (int k = 0; k < L; ++k) {
    const R hxx = in.mh[k], hzz = in.mh[L + k], hxz = in.mh[2 * L + k];
    exh = exh + dt * d.a1[k] * hxx;
    ezh_from_M = ezh_from_M + dt * d.a1[k] * hzz;
    gh = gh + dt * d.a1[k] * hxz;
    Mzzh_sum = Mzzh_sum + d.a1[k] * hzz;
    nm[k] = (R(1) + dt * d.a2[k]) * hxx + SMxxh;
    nm[L + k] = (R(1) + dt * d.a2[k]) * hzz + SMzzh;
    nm[2 * L + k] = (R(1) + dt * d.a2[k]) * hxz + SMxzh;
  }
  ezh = ezh + ezh_from_M;
  if (HERM && fs_row) ezh = ezh + esh;  // es = ezz* of this row feeds row 5 (formula #3)
  R ezz_raw_h = ezh;
  if (fs_row) {  // ezz* = -(C13_0 exx + C13_1 SMxx + C33_1 SMzz)/C33_0 replaces the raw rate
    exh = exh - l0 / pz0 * ezh;
    fsMxx = -l1 / pz0 * ezh;
    fsMzz = -pz1 / pz0 * ezh;
    for (int k = 0; k < L; ++k) {
      nm[k] = nm[k] + fsMxx;
      nm[L + k] = nm[L + k] + fsMzz;
    }
    ezz_raw_h = 0;
  }
  for (int k = 0; k < 3 * L; ++k) gst(Ma[k], cb, nm[k]);
  // gradient with respect to C11, C13, C33 (lam and mu through them when isotropic),
  // C55 at the sxz node and tau of this cell
  {
    const R tt = R(1) + t;
    R g11, g13, g33, gtau;
    if (!fs_row) {
      const R ax_ = dt * (tt * exx + t * SMxx), az_ = dt * (tt * ezz + t * SMzz);
      g11 = sxxh * ax_;
      g33 = szzh * az_;
      g13 = sxxh * az_ + szzh * ax_;
      gtau = sxxh * dt * (px * exx + lr * ezz + px * SMxx + lr * SMzz) +
             szzh * dt * (lr * exx + pz * ezz + lr * SMxx + pz * SMzz);
    } else {
      // sxx' = sxx + dt (A + C13_0 z) with z = -N/C33_0; Mzz'_n depends on z via dt a1_n z
      const R z = ezs;
      const R q = N / (pz0 * pz0);
      const R dz13 = -(tt * exx + t * SMxx) / pz0;
      const R dz33 = -(t * SMzz) / pz0 + q * tt;
      const R dzt = -(lr * exx + lr * SMxx + pz * SMzz) / pz0 + q * pz;
      g11 = sxxh * dt * (tt * exx + t * SMxx);
      g13 = sxxh * dt * (t * SMzz + tt * z + l0 * dz13) + dt * Mzzh_sum * dz13;
      g33 = sxxh * dt * (l0 * dz33) + dt * Mzzh_sum * dz33;
      gtau = sxxh * dt * (px * exx + px * SMxx + lr * SMzz + lr * z + l0 * dzt) +
             dt * Mzzh_sum * dzt;
      if (HERM) {  // z also enters row 5's d vz/dz
        g13 = g13 + esh * dz13;
        g33 = g33 + esh * dz33;
        gtau = gtau + esh * dzt;
      }
    }
    gtau = gtau + sxzh * dt * (mxz * gxz + mxz * SMxz);
    if (d.vti) {
      gst(d.g_lam, cb, in.acc_lam + g13);
      gst(d.g_mu, cb, in.acc_mu + g11);
      gst(d.g_c33, cb, in.acc_c33 + g33);
    } else {  // C11 = C33 = lam + 2 mu, C13 = lam
      gst(d.g_lam, cb, in.acc_lam + ((g11 + g33) + g13));
      gst(d.g_mu, cb, in.acc_mu + two * (g11 + g33));
    }
    gst(d.g_muxz, cb, in.acc_muxz + sxzh * dt * (tt * gxz + t * SMxz));
    if (L > 0) gst(d.g_tau, cb, in.acc_tau + gtau);
  }
  // CPML adjoint and the adjoints of the raw derivatives
  R exx_r = exh, exz_r = gh, ezx_r = gh, ezz_r = ezz_raw_h;
  if (xk >= 0) {
    exx_r = cpml_adj(exh, &p0h, axv, bxv, kxv);
    exz_r = cpml_adj(gh, &p1h, axv, bxv, kxv);
    gst(d.apsi[0], qx, p0h);
    gst(d.apsi[1], qx, p1h);
  }
  if (zk >= 0) {
    ezx_r = cpml_adj(gh, &p4h, azv, bzv, kzv);
    ezz_r = cpml_adj(ezz_raw_h, &p5h, azv, bzv, kzv);
    gst(d.apsi[4], qz, p4h);
    gst(d.apsi[5], qz, p5h);
  }
  gst(d.ehat[0], cb, exx_r);
  gst(d.ehat[1], cb, ezx_r);
  gst(d.ehat[2], cb, ezz_r);
  gst(d.ehat[3], cb, exz_r);
  return ezz_r;
}

// ---- adjoint of the velocity half step (pointwise part of one interior cell) ----------
// dxx = D+x sxx, dxzz = D-z sxz, dxzx = D-x sxz, dzz = D+z szz: the raw forward derivatives;
// vxh, vzh: the velocity adjoint of the cell
// Everything one cell of the transposed velocity half reads from memory (see AdjStressIn),
// including the velocity adjoint that the gather updates.
template <typename R>
struct AdjVelIn {
  R axv = 0, bxv = 0, kxv = 1, azv = 0, bzv = 0, kzv = 1;
  R p2h = 0, p3h = 0, p6h = 0, p7h = 0;
  R fx, fz;  // the forward step's stress divergence (stretched, sources included)
  R rho, rho1, rho2, acc_bx, acc_bz, avx, avz;
  __device__ __forceinline__ void load(const AdjDesc<R>& d, int i, int j) {
    constexpr unsigned E = sizeof(R);
    const int ldx = d.ldx;
    const unsigned cb = (unsigned)(j * ldx + i) * E;
    const unsigned mb = (unsigned)(j * d.ldm + i) * E;
    const int xk = d.T ? strip_x(i, d.nx, d.T, d.xlo, d.xhi) : -1;
    const int zk = d.T ? strip_x(j, d.nz, d.T, d.zlo, d.zhi) : -1;
    const unsigned qx = (unsigned)(j * 2 * d.T + xk) * E, qz = (unsigned)(zk * ldx + i) * E;
    if (xk >= 0) {
      axv = gld(d.ax, i * E); bxv = gld(d.bx, i * E); kxv = gld(d.kx, i * E);
      p2h = gld((const R*)d.apsi[2], qx); p3h = gld((const R*)d.apsi[3], qx);
    }
    if (zk >= 0) {
      azv = d.az[j]; bzv = d.bz[j]; kzv = d.kz[j];
      p6h = gld((const R*)d.apsi[6], qz); p7h = gld((const R*)d.apsi[7], qz);
    }
    fx = gld_nt(d.ffx, cb);
    fz = gld_nt(d.ffz, cb);
    rho = gld(d.rho, mb);
    rho1 = gld(d.rho, mb + E);
    rho2 = gld(d.rho, mb + d.ldm * E);
    acc_bx = gld((const R*)d.g_bx, cb);
    acc_bz = gld((const R*)d.g_bz, cb);
    avx = gld((const R*)d.avx, cb);
    avz = gld((const R*)d.avz, cb);
  }
};

template <typename R>
__device__ __forceinline__ void adj_vel_cell(const AdjDesc<R>& d, const AdjVelIn<R>& in, int i,
                                             int j, R vxh, R vzh) {
  constexpr unsigned E = sizeof(R);
  const int ldx = d.ldx;
  const unsigned cb = (unsigned)(j * ldx + i) * E;
  const R dt = d.dt, two = R(2);
  const int xk = d.T ? strip_x(i, d.nx, d.T, d.xlo, d.xhi) : -1;
  const int zk = d.T ? strip_x(j, d.nz, d.T, d.zlo, d.zhi) : -1;
  const unsigned qx = (unsigned)(j * 2 * d.T + xk) * E, qz = (unsigned)(zk * ldx + i) * E;
  const R axv = in.axv, bxv = in.bxv, kxv = in.kxv, azv = in.azv, bzv = in.bzv, kzv = in.kzv;
  R p2h = in.p2h, p3h = in.p3h, p6h = in.p6h, p7h = in.p7h;
  const R rho = in.rho, rho1 = in.rho1, rho2 = in.rho2;
  const R acc_bx = in.acc_bx, acc_bz = in.acc_bz;
  const R fx = in.fx, fz = in.fz;  // stored by the recording kernel
  const R bxb = two / (rho + rho1);
  const R bzb = two / (rho + rho2);
  // gradient with respect to the buoyancies (scattered to rho at the end)
  gst(d.g_bx, cb, acc_bx + vxh * dt * fx);
  gst(d.g_bz, cb, acc_bz + vzh * dt * fz);
  const R hx = dt * bxb * vxh, hz = dt * bzb * vzh;
  R hxx = hx, hxzz = hx, hzz = hz, hxzx = hz;
  if (xk >= 0) {
    hxx = cpml_adj(hx, &p2h, axv, bxv, kxv);
    hxzx = cpml_adj(hz, &p3h, axv, bxv, kxv);
    gst(d.apsi[2], qx, p2h);
    gst(d.apsi[3], qx, p3h);
  }
  if (zk >= 0) {
    hxzz = cpml_adj(hx, &p6h, azv, bzv, kzv);
    hzz = cpml_adj(hz, &p7h, azv, bzv, kzv);
    gst(d.apsi[6], qz, p6h);
    gst(d.apsi[7], qz, p7h);
  }
  gst(d.dhat[0], cb, hxx);
  gst(d.dhat[1], cb, hxzz);
  gst(d.dhat[2], cb, hxzx);
  gst(d.dhat[3], cb, hzz);
}

// (G_a - G_c)^T e_A at target row tg (4..11, wave-uniform): sum over the nrows source rows of
// table rows t0 .. t0+nrows-1; q = LDS address of the first source row, st = LDS row pitch.
template <typename R>
__device__ __forceinline__ R afda_corr(const R* q, int st, const R* tab, int t0, int nrows,
                                       int tg) {
  const __attribute__((address_space(4))) R* c =
      (const __attribute__((address_space(4))) R*)tab + 104 + t0 * 8 +
      __builtin_amdgcn_readfirstlane(tg - 4);
  R v = c[0] * q[0];
  for (int k = 1; k < nrows; ++k) v = v + c[k * 8] * q[k * st];
  return v;
}

// ---- tiled transposed steps ------------------------------------------------------------
// The adjoint kernels use the forward kernels' tiles: 64 columns x (4 RZ) rows, 4 wavefronts,
// XCD-aware tile order, every stenciled field staged once in LDS with its halo (TileLoad).
// The sums are the ones of the one-cell formulas, in the same order.

// u_hat += R^T r^n for the velocity receivers (k_adj_inject's sums, in the same order), by the
// threads of a transposed-stress launch: no other thread of it reads or writes v_hat, and
// k_adj_vel runs after it.
template <typename R>
__device__ __forceinline__ void inject_velocity_targets(const AdjDesc<R>& d) {
  const int lane = threadIdx.x, w = threadIdx.y;
  for (int t = (int)blockIdx.x * (TX * NW) + w * TX + lane; t < d.ntgt_v;
       t += (int)gridDim.x * (TX * NW)) {
    R s = 0;
    for (int k = d.tptr[t]; k < d.tptr[t + 1]; ++k) s = s + d.resid_row[d.trec[k]];
    R* p = (d.tcomp[t] == 0 ? d.avx : d.avz) + d.tcell[t];
    *p = *p + s;
  }
}

// The pointwise transposed stress half of a wave's RZ rows (loaded in `in`). Formula #3
// (fs = 3): row 5's d vz/dz reads the surface row's ezz*, so the surface row's cell needs row
// 5's raw-rate adjoint: row 5 goes first and hands it over through LDS (the first tile row
// holds rows 4..4+TZ-1; the condition is uniform over the workgroup, and every wave of the
// workgroup reaches the barrier).
template <typename R, int L, int RZ, bool HERM>
__device__ __forceinline__ void adj_stress_rows(const AdjDesc<R>& d, const AdjStressIn<R, L>* in,
                                                int i, int j0, bool col) {
  const int lane = threadIdx.x, w = threadIdx.y;
  if constexpr (HERM) {
    __shared__ R s_ez[TX];
    const bool herm = j0 == 4 && 5 < d.nz - 4;
    if (herm) {
#pragma unroll
      for (int r = 0; r < RZ; ++r)
        if (col && j0 + w * RZ + r == 5)
          s_ez[lane] = adj_stress_cell<R, L, true>(d, in[r], i, 5);
      __syncthreads();
    }
    if (!col) return;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int j = j0 + w * RZ + r;
      if (j >= d.nz - 4) return;
      if (herm && j == 5) continue;
      R esh = R(0);
      if (herm && j == 4)
        esh = ((const __attribute__((address_space(4))) R*)d.afda)[216] * s_ez[lane];  // alpha
      adj_stress_cell<R, L, true>(d, in[r], i, j, esh);
    }
  } else {
    if (!col) return;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
      const int j = j0 + w * RZ + r;
      if (j >= d.nz - 4) return;
      adj_stress_cell<R, L, false>(d, in[r], i, j);
    }
  }
}

// Transposed stress half, pointwise: every forward quantity is stored per cell, so the tile
// needs no LDS (but for formula #3's hand-over); the tile shape keeps the loads coalesced.
// HERM: the variant for formula #3 (fs = 3); the others carry none of its code.
template <typename R, int L, int RZ, bool HERM>
__global__ __launch_bounds__(TX* NW) void k_adj_stress(const AdjDesc<R> d) {
  constexpr int TZ = NW * RZ;
  const int lane = threadIdx.x, w = threadIdx.y;
  inject_velocity_targets(d);
  const int ntx = (d.nx - 8 + TX - 1) / TX;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i0 = 4 + (tile % ntx) * TX, j0 = 4 + (tile / ntx) * TZ;
  const int i = i0 + lane;
  const bool col = i < d.nx - 4;
  if (!HERM && !col) return;  // (the formula #3 variant's lanes all reach its barrier)
  AdjStressIn<R, L> in[RZ];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + w * RZ + r;
    if (col && j < d.nz - 4) in[r].load(d, i, j);
  }
  adj_stress_rows<R, L, RZ, HERM>(d, in, i, j0, col);
}

// Transposed stress half, gather (v_hat -= D^T e_hat), fused with the transposed velocity
// half, pointwise (which needs only the cell's new v_hat and the stored stress divergence):
// e_hat from LDS.
template <typename R, int RZ>
__global__ __launch_bounds__(TX* NW) void k_adj_vel(const AdjDesc<R> d) {
  constexpr int TZ = NW * RZ, WX = TX + 8;
  constexpr unsigned E = sizeof(R);
  __shared__ R s_e0[TZ * WX];        // e_hat exx: x halo
  __shared__ R s_e3[TZ * WX];        // e_hat exz: x halo
  __shared__ R s_e1[(TZ + 8) * TX];  // e_hat ezx: z halo
  __shared__ R s_e2[(TZ + 8) * TX];  // e_hat ezz: z halo
  const int lane = threadIdx.x, w = threadIdx.y;
  const int ntx = (d.nx - 8 + TX - 1) / TX;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i0 = 4 + (tile % ntx) * TX, j0 = 4 + (tile / ntx) * TZ;
  const int ldx = d.ldx;
  TileLoad<R, TZ, 4> t_e0, t_e3;
  TileLoad<R, TZ + 8, 0> t_e1, t_e2;
  t_e0.fetch(d.ehat[0], j0, i0, ldx, lane, w);
  t_e3.fetch(d.ehat[3], j0, i0, ldx, lane, w);
  t_e1.fetch(d.ehat[1], j0 - 4, i0, ldx, lane, w);
  t_e2.fetch(d.ehat[2], j0 - 4, i0, ldx, lane, w);
  const int i = i0 + lane;
  AdjVelIn<R> in[RZ];
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + w * RZ + r;
    if (i < d.nx - 4 && j < d.nz - 4) in[r].load(d, i, j);
  }
  t_e0.store(s_e0, lane, w);
  t_e3.store(s_e3, lane, w);
  t_e1.store(s_e1, lane, w);
  t_e2.store(s_e2, lane, w);
  __syncthreads();
  if (i >= d.nx - 4) return;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + w * RZ + r;
    if (j >= d.nz - 4) return;
    const int tr = w * RZ + r;
    const unsigned cb = (unsigned)(j * ldx + i) * E;
    // exx = D-x vx, ezx = D+z vx, ezz = D-z vz, exz = D+x vz
    R gvx = op_p(&s_e0[tr * WX + lane + 4], 1, d.Dx) +
            op_m(&s_e1[(tr + 4) * TX + lane], TX, d.Dz);
    R gvz = op_p(&s_e2[(tr + 4) * TX + lane], TX, d.Dz) +
            op_m(&s_e3[tr * WX + lane + 4], 1, d.Dx);
    const int ju = __builtin_amdgcn_readfirstlane(j0 + w * RZ) + r;  // j, wave-uniform
    if (d.fs >= 2 && ju < 12) {  // H-AFDA: rows 4..7 hold the source rows (LDS row 8 - j0 + k)
      const int s0 = 8 - j0;
      gvx = gvx - afda_corr(&s_e1[s0 * TX + lane], TX, d.afda, 10, 3, ju);      // D0..2
      gvz = gvz - afda_corr(&s_e2[(s0 + 1) * TX + lane], TX, d.afda, 7, 3, ju); // C1..3
    }
    const R vxh = in[r].avx - gvx;
    const R vzh = in[r].avz - gvz;
    gst(d.avx, cb, vxh);
    gst(d.avz, cb, vzh);
    adj_vel_cell<R>(d, in[r], i, j, vxh, vzh);
  }
}

// Transposed velocity half, gather: s_hat -= D^T d_hat over rows [row0, nz-4). Rows 0..3 are
// the free-surface image rows (row0 = 0 with imaging), whose adjoint the next k_adj_stress
// moves to the interior rows they copy. Their z-stencils reach rows -4..-1, so the d_hat planes are read
// from a base moved back by 4 rows into the zero guard rows in front of each plane: every
// 32-bit byte offset stays non-negative.
template <typename R, int RZ>
__global__ __launch_bounds__(TX* NW) void k_adj_sgather(const AdjDesc<R> d, int row0) {
  constexpr int TZ = NW * RZ, WX = TX + 8;
  constexpr unsigned E = sizeof(R);
  __shared__ R s_d0[TZ * WX];        // d_hat dxx = D+x sxx: x halo
  __shared__ R s_d2[TZ * WX];        // d_hat dxzx = D-x sxz: x halo
  __shared__ R s_d1[(TZ + 8) * TX];  // d_hat dxzz = D-z sxz: z halo
  __shared__ R s_d3[(TZ + 8) * TX];  // d_hat dzz = D+z szz: z halo
  const int lane = threadIdx.x, w = threadIdx.y;
  const int ntx = (d.nx - 8 + TX - 1) / TX;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i0 = 4 + (tile % ntx) * TX, j0 = row0 + (tile / ntx) * TZ;
  const int ldx = d.ldx;
  const size_t back = (size_t)4 * ldx;  // guarded bases: row r of the plane is row r + 4
  TileLoad<R, TZ, 4> t_d0, t_d2;
  TileLoad<R, TZ + 8, 0> t_d1, t_d3;
  t_d0.fetch(d.dhat[0] - back, j0 + 4, i0, ldx, lane, w);
  t_d2.fetch(d.dhat[2] - back, j0 + 4, i0, ldx, lane, w);
  t_d1.fetch(d.dhat[1] - back, j0, i0, ldx, lane, w);
  t_d3.fetch(d.dhat[3] - back, j0, i0, ldx, lane, w);
  const int i = i0 + lane;
  R oxx[RZ], oxz[RZ], ozz[RZ];  // the stress adjoint of the wave's cells, before the barrier
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + w * RZ + r;
    if (i < d.nx - 4 && j < d.nz - 4) {
      const unsigned cb = (unsigned)(j * ldx + i) * E;
      oxx[r] = gld((const R*)d.asxx, cb);
      oxz[r] = gld((const R*)d.asxz, cb);
      ozz[r] = gld((const R*)d.aszz, cb);
    }
  }
  t_d0.store(s_d0, lane, w);
  t_d2.store(s_d2, lane, w);
  t_d1.store(s_d1, lane, w);
  t_d3.store(s_d3, lane, w);
  __syncthreads();
  if (i >= d.nx - 4) return;
#pragma unroll
  for (int r = 0; r < RZ; ++r) {
    const int j = j0 + w * RZ + r;
    if (j >= d.nz - 4) return;
    const int tr = w * RZ + r;
    const unsigned cb = (unsigned)(j * ldx + i) * E;
    const R gsxx = op_m(&s_d0[tr * WX + lane + 4], 1, d.Dx);
    R gsxz = op_p(&s_d1[(tr + 4) * TX + lane], TX, d.Dz) +
             op_p(&s_d2[tr * WX + lane + 4], 1, d.Dx);
    R gszz = op_m(&s_d3[(tr + 4) * TX + lane], TX, d.Dz);
    const int ju = __builtin_amdgcn_readfirstlane(j0 + w * RZ) + r;  // j, wave-uniform
    if (d.fs >= 2 && ju < 12) {  // H-AFDA (row0 = 4): source rows 4..7 at LDS row 8 - j0 + k
      const int s0 = 8 - j0;
      gsxz = gsxz - afda_corr(&s_d1[s0 * TX + lane], TX, d.afda, 0, 4, ju);  // A0..3
      gszz = gszz - afda_corr(&s_d3[s0 * TX + lane], TX, d.afda, 4, 3, ju);  // B0..2
    }
    gst(d.asxx, cb, oxx[r] - gsxx);
    gst(d.asxz, cb, oxz[r] - gsxz);
    gst(d.aszz, cb, ozz[r] - gszz);
  }
}

// ---- residual injection: u_hat += R^T r^n (receivers sample after the stress half) ----
// One thread per distinct (component, cell) target; receivers that share a target are
// summed in ascending receiver order (tptr/trec, a CSR list), so the result does not depend
// on thread scheduling.
template <typename R>
__global__ __launch_bounds__(256) void k_adj_inject(R* f0, R* f1, R* f2, R* f3, R* f4,
                                                    const int* tcell, const int* tcomp,
                                                    const int* tptr, const int* trec, int ntgt,
                                                    const R* resid_row) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntgt) return;
  R s = 0;
  for (int k = tptr[t]; k < tptr[t + 1]; ++k) s = s + resid_row[trec[k]];
  R* f;
  switch (tcomp[t]) {
    case 0: f = f0; break;
    case 1: f = f1; break;
    case 2: f = f2; break;
    case 3: f = f3; break;
    default: f = f4; break;
  }
  R* p = f + tcell[t];
  *p = *p + s;
}

// ---- residual and misfit: r = d - d_obs with the traces and the residual [nt][nrec] and
// the observed data [nrec][nt]; J = 1/2 sum r^2 accumulated in double, per receiver by one
// workgroup (256 strided partial sums over time, then a fixed pairwise tree) and then over
// the receivers in order (a fixed order: the same J on every run)
template <typename R>
__global__ __launch_bounds__(256) void k_residual(const R* __restrict__ traces,
                                                  const R* __restrict__ dobs,
                                                  R* __restrict__ resid, int nt, int nrec) {
  const size_t n = (size_t)nt * nrec;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t t = k / (size_t)nrec, r = k % (size_t)nrec;
    resid[k] = traces[k] - dobs[r * nt + t];
  }
}
// one workgroup of 256 threads per receiver (grid = nrec)
template <typename R>
__global__ __launch_bounds__(256) void k_misfit_rec(const R* __restrict__ resid, int nt, int nrec,
                                                    double* __restrict__ part) {
  __shared__ double s_part[256];
  const int r = blockIdx.x, k = threadIdx.x;
  double s = 0.0;
  for (int t = k; t < nt; t += 256) {
    const double x = (double)resid[(size_t)t * nrec + r];
    s += 0.5 * x * x;
  }
  s_part[k] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (k < h) s_part[k] += s_part[k + h];
    __syncthreads();
  }
  if (k == 0) part[r] = s_part[0];
}
__global__ void k_misfit_sum(const double* __restrict__ part, int nrec, double* __restrict__ J) {
  if (blockIdx.x * blockDim.x + threadIdx.x != 0) return;
  double s = 0.0;
  for (int r = 0; r < nrec; ++r) s += part[r];
  *J = s;
}

// ---- batched device-to-device copy: up to 16 regions per launch (blockIdx.y = region).
// Regions start 16-byte aligned and hold a multiple of 4 bytes.
struct CopyList {
  const char* src[16];
  char* dst[16];
  unsigned long long bytes[16];
  int n;
};

__global__ __launch_bounds__(256) void k_copy_list(const CopyList c) {
  const int e = blockIdx.y;
  if (e >= c.n) return;
  const char* s = c.src[e];
  char* d = c.dst[e];
  const size_t nb = c.bytes[e];
  const size_t n16 = nb / 16;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = tid; k < n16; k += stride) ((uint4*)d)[k] = ((const uint4*)s)[k];
  const size_t t0 = n16 * 16;
  if (tid < 4 && t0 + 4 * tid < nb)
    ((unsigned*)(d + t0))[tid] = ((const unsigned*)(s + t0))[tid];
}

// ---- gradient finalisation: mu_xz and buoyancy accumulators to mu and rho --------------
template <typename R>
__global__ __launch_bounds__(256) void k_grad_finalize(const AdjDesc<R> d, R* g_rho_out,
                                                       R* g_mu_out) {
  constexpr unsigned E = sizeof(R);
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx || j >= d.nz) return;
  const unsigned cb = (unsigned)(j * d.ldx + i) * E;
  const unsigned mb = (unsigned)(j * d.ldm + i) * E;
  const int rb = d.ldx * (int)E;
  const int mrb = d.ldm * (int)E;
  const R quarter = R(0.25), two = R(2);
  // mu_xz(c') = 1/4 (mu(c') + mu(c'+1) + mu(c'+ld) + mu(c'+ld+1)): mu(c) feeds c, c-1,
  // c-ld, c-ld-1
  R gm_ = d.vti ? R(0) : gld((const R*)d.g_mu, cb);  // VTI: C55 acts only at sxz nodes
  R gx = gld((const R*)d.g_muxz, cb);
  if (i >= 1) gx = gx + gld((const R*)d.g_muxz, cb - E);
  if (j >= 1) gx = gx + gld((const R*)d.g_muxz, cb - rb);
  if (i >= 1 && j >= 1) gx = gx + gld((const R*)d.g_muxz, cb - rb - E);
  gst(g_mu_out, cb, gm_ + quarter * gx);
  // bx(c) = 2/(rho(c) + rho(c+1)), bz(c) = 2/(rho(c) + rho(c+ld)); d b / d rho = -b^2/2
  auto bxat = [&](unsigned m) { return two / (gld(d.rho, m) + gld(d.rho, m + E)); };
  auto bzat = [&](unsigned m) { return two / (gld(d.rho, m) + gld(d.rho, m + mrb)); };
  R gr = 0;
  R b = bxat(mb);
  gr = gr - gld((const R*)d.g_bx, cb) * b * b / two;
  b = bzat(mb);
  gr = gr - gld((const R*)d.g_bz, cb) * b * b / two;
  if (i >= 1) {
    b = bxat(mb - E);
    gr = gr - gld((const R*)d.g_bx, cb - E) * b * b / two;
  }
  if (j >= 1) {
    b = bzat(mb - mrb);
    gr = gr - gld((const R*)d.g_bz, cb - rb) * b * b / two;
  }
  gst(g_rho_out, cb, gr);
}

}  // namespace vi
