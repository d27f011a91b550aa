// Fused substep path: the arithmetic of the per-routine kernels in
// fg_steps.hip (same expressions, same rounding), regrouped so that one
// substep makes five passes over HBM instead of ~100:
//
//   k_src   2D tile (TH x TW cells + halo) staged in LDS:
//           CorrectVtheta, evanescent damping, pressure, potential,
//           Substep1, viscous stress, viscous update, artificial
//           viscosity, velocity update, second damping pass
//           state A -> state B
//   k_trad  column-marching radial van Leer sweep of the five transported
//           quantities (each slope and flux evaluated once), plus one wave
//           per ring for VMed (sequential ring sum)
//           state B -> R (Sigma, RadMomP/M, ThetaMomP/M)
//   k_tazw  one wave per contiguous chunk of a ring: FARGO residual and
//           integer shift, both azimuthal sweeps, the shift, vtheta
//           recovery; R -> X (Sigma, vtheta, RadMomP, RadMomM)
//   k_fin   one ring per workgroup: vrad recovery, third damping pass,
//           temperature, and the next substep's CFL ring minimum
//           X -> state A
//   force   DeviceReduce-order disk force on the star (partials as in
//           fg_reduce.hip) and the CFL count in one final kernel
//           (k_force_cfl_final; FP32 mode: k_force_rings_cfl)
//
// Adiabatic decks add Substep3 (k_src3) between k_src and k_trad and carry
// the energy through the sweeps and k_fin; self-gravitating decks run the
// SG field (fg_sg_fused) first and add its kick in k_src.
//
// Supported decks: evanescent boundaries, per-substep CFL; isothermal or
// adiabatic, with or without self-gravity (FP32 bulk mode: isothermal
// without self-gravity).  Everything else uses the per-routine path.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fg_cell.h"
#include "fg_internal.h"
#include "fg_reduce.h"

#define TH 16
#define TW 64
#define HALO_R 4  // rows above the tile (i - 4 .. )
#define HALO_B 2  // rows below the tile ( .. i + TH + 1)
#define HALO_C 4  // columns on each side
#define HC (TW + 2 * HALO_C)
// sector of a staged column j0 - HALO_C + c (j0 = the tile's first column):
// at most one ring length outside [0, ns) once ns reaches the staged width,
// so a compare and an add or subtract replace the two runtime modulos
__device__ __forceinline__ int wrap_col(int x, int ns) {
  if (x < 0) x += ns;
  while (x >= ns) x -= ns;
  return x;
}
#define SRC64_TPB 512  // k_src: two 80 KB workgroups per CU, four waves per SIMD

fg_status fg_disk_force_async(fg_ctx *c, const double *dens, double x, double y, double rs, double *out_dev);



struct fused_bufs {
  double *Bd, *Bvr, *Bvt;                // after k_src
  double *Rd, *Rrmp, *Rrmm, *Rtmp, *Rtmm;  // after k_trad
  double *Xd, *Xvt, *Xrmp, *Xrmm;        // after k_taz
  double *P;                             // pressure (temperature output only)
  // adiabatic decks: EnergyInt after k_src, energy after k_trad and k_taz;
  // self-gravity: the kick of VradInt / VthetaInt (NULL when not allocated)
  double *Beint, *Re, *Xe, *kick_r, *kick_t;
  // FP32 bulk mode: B, R and X stored in FP32 in the same arrays
  float *fBd() const { return (float *)Bd; }
  float *fBvr() const { return (float *)Bvr; }
  float *fBvt() const { return (float *)Bvt; }
  float *fR(int q) const { return (float *)(q == 0 ? Rrmp : q == 1 ? Rrmm : q == 2 ? Rtmp : q == 3 ? Rtmm : Rd); }
  float *fXd() const { return (float *)Xd; }
  float *fXvt() const { return (float *)Xvt; }
  float *fXrmp() const { return (float *)Xrmp; }
  float *fXrmm() const { return (float *)Xrmm; }
};

struct src_args {
  const double *dens, *vr, *vt;
  double *odens, *ovr, *ovt, *opress;
  const double *nu;   // viscosity of this substep (viscous stress, second damping pass)
  const double *nu1;  // viscosity as the first damping pass sees it (an adiabatic
                      // run refreshes it after that pass on a DT's first substep)
  // adiabatic: energy in, EnergyInt out (SourceEuler.cu:281, Kernels.cu:474-477)
  const double *energy;
  double *oeint;
  double gamma;
  // self-gravity: the kick of Update_sgvelocityKernel per cell (Kernels.cu:915-932,
  // formed by k_sg_acc_fused), and the ring means of the radial SG acceleration
  // that the first (sg1: previous evaluation) and second (sg2: this substep's)
  // damping passes take for vtheta0
  const double *kick_r, *kick_t;
  const double *sg1, *sg2;
  fg_dev_tables t;
  fg_planets_arg P;
  double dt, omega_frame, domega;
  int nr, ns, indirect, evanescent, corotating;
};

// potential at one cell (FillForcesArraysKernel, Kernels.cu:1234-1252)
// InvDistance = 1.0 / Rm is passed in: the ring table holds the same
// correctly rounded quotient
__device__ __forceinline__ double cell_potential(const fg_planets_arg &P, double Rm, double InvDistance, double cosp,
                                                 double sinp, int indirect) {
  const double x = Rm * cosp, y = Rm * sinp;
  double acc = 0.0;
  for (int k = 0; k < P.nb; k++) {
    double xp = P.x[k], yp = P.y[k], m = P.m[k];
    double distance = (x - xp) * (x - xp) + (y - yp) * (y - yp);
    double distancesmooth = sqrt(distance + P.smooth[k]);
    double pv = -1.0 * m / distancesmooth;
    if (indirect) pv += 1.0 * m * P.invd3[k] * (x * xp + y * yp);
    acc += pv;
    if (k == 0) {
      pv = -1.0 * 1.0 * InvDistance;
      pv -= P.ix * x + P.iy * y;
      acc += pv;
    }
  }
  return acc;
}

// evanescent damping constants of ring i for a step (Kernels.cu:978-1004);
// with self-gravity vtheta0 takes the ring mean of the radial SG
// acceleration (sgacc, NULL without self-gravity)
__device__ __forceinline__ void ev_consts(const fg_dev_tables &t, const double *nu, int i, double step,
                                          double omega_frame, double &lambda, double &vrad0, double &vtheta0,
                                          const double *sgacc = nullptr) {
  const double Rm = t.Rmed[i];
  lambda = t.ev_coef[i] * step / t.ev_period[i];
  vtheta0 = sgacc ? sqrt(t.ev_vtheta_arg[i] - Rm * sgacc[i]) : sqrt(t.ev_vtheta_arg[i]);
  vtheta0 -= Rm * omega_frame;
  vrad0 = -3.0 * nu[i] / Rm * t.vrad0_slope;
}

// per-ring tables the source stages read, staged per tile row in LDS
enum { RT_MED, RT_IMED, RT_IDMED, RT_INF, RT_IINF, RT_IDXT, RT_SUPP, RT_IDSUP, RT_SUP, RT_NU, RT_CS, RT_SMED, RT_EVZ, RT_EMED, RT_N };
__device__ __forceinline__ const double *ring_table(const fg_dev_tables &t, const double *nu, int x) {
  switch (x) {
    case RT_MED: return t.Rmed;
    case RT_IMED: return t.invRmed;
    case RT_IDMED: return t.invdiffRmed;
    case RT_INF: return t.Rinf;
    case RT_IINF: return t.invRinf;
    case RT_IDXT: return t.invdxtheta;
    case RT_SUPP: return t.supp_torque;
    case RT_IDSUP: return t.invdiffRsup;
    case RT_SUP: return t.Rsup;
    case RT_NU: return nu;
    case RT_CS: return t.cs_iso;
    default: return t.SigmaMed;
  }
}

// MODE bit 1: adiabatic (energy damped, P = (gamma - 1) e, EnergyInt out);
// bit 2: self-gravity kick after Substep1
template <int KTH, int KTW, int KTPB, int MODE>
__global__ __launch_bounds__(KTPB) void k_src(src_args a) {
  constexpr bool ADI = MODE & 1, SGK = MODE & 2;
  constexpr int KHR = KTH + HALO_R + HALO_B, KHC = KTW + 2 * HALO_C;
  constexpr int KPT = (KHR * KHC + KTPB - 1) / KTPB;
  __shared__ double sS[KHR][KHC];   // Sigma after the first damping
  __shared__ double sA[KHR][KHC];   // pressure, then VradInt
  __shared__ double sB[KHR][KHC];   // potential, then VthetaInt
  __shared__ double sC[KHR][KHC];   // vtheta, then TAURR, then Qr
  __shared__ double sD[KHR][KHC];   // vrad, then TAUPP, then Qtheta
  __shared__ double sE[KHR][KHC];   // TAURP
  const int nr = a.nr, ns = a.ns;
  const int i0 = blockIdx.y * KTH, j0 = blockIdx.x * KTW;
  const int tid = threadIdx.x;
  const fg_dev_tables &t = a.t;
  const double dt = a.dt;
#define GI(r) (i0 - HALO_R + (r))
#define GJ(c) wrap_col(j0 - HALO_C + (c), ns)

  // ---- S0: load, CorrectVtheta, damping #1, pressure, potential ----
  // all of a thread's loads are issued before any of them is used
  double ls[KPT], lvr[KPT], lvt[KPT], le[ADI ? KPT : 1], lkr[SGK ? KPT : 1], lkt[SGK ? KPT : 1];
#pragma unroll
  for (int k = 0; k < KPT; k++) {
    const int q = tid + k * KTPB;
    ls[k] = lvr[k] = lvt[k] = 0.0;
    if constexpr (ADI) le[k] = 0.0;
    if constexpr (SGK) lkr[k] = lkt[k] = 0.0;
    if (q < KHR * KHC) {
      const int r = q / KHC, cc = q - r * KHC, i = GI(r);
      if (i >= 0 && i <= nr) {
        const int j = GJ(cc);
        const long l = (long)i * ns + j;
        ls[k] = a.dens[l];
        lvr[k] = a.vr[l];
        lvt[k] = a.vt[l];
        if constexpr (ADI) le[k] = a.energy[l];
        if constexpr (SGK)
          if (i < nr) {
            lkr[k] = a.kick_r[l];
            lkt[k] = a.kick_t[l];
          }
      }
    }
  }
  // The per-ring tables of the tile's rows (rings clamped into the table;
  // rows outside the disk are never read) and the potential's cos/sin of
  // its columns, staged in LDS while the loads above are in flight: every
  // stage below reads them from there.  Read from memory inside a stage,
  // each would be a dependent round trip.  Each wave loads every fourth
  // table, lane r row r.
  __shared__ double rt[RT_N][KHR];
  __shared__ double ct[2][KHC];
  {
    const int w = tid >> 6, ln = tid & 63;
    if (w < 4 && ln < KHR) {
      const int i = min(max(GI(ln), 0), nr - 1);
      // wave w: tables w, w + 4, w + 8 and (wave 0) the ev_zone flags;
      // all loads issued before the first LDS write
      const double v0 = ring_table(t, a.nu, w)[i], v1 = ring_table(t, a.nu, w + 4)[i];
      const double v2 = ring_table(t, a.nu, w + 8)[i];
      const int ez = t.ev_zone[i];
      rt[w][ln] = v0;
      rt[w + 4][ln] = v1;
      rt[w + 8][ln] = v2;
      if (w == 0) rt[RT_EVZ][ln] = (double)ez;
      if (ADI && w == 1) rt[RT_EMED][ln] = t.EnergyMed[i];
    }
    for (int c = tid; c < KHC; c += KTPB) {
      const int j = GJ(c);
      ct[0][c] = t.cos_pot[j];
      ct[1][c] = t.sin_pot[j];
    }
  }
  // damping constants of the tile's rings, once per ring (Kernels.cu:978-1004):
  // lambda, then vrad0 and vtheta0 of the first pass, then of the second
  __shared__ double evc[KHR][5];
  if (a.evanescent) {
    for (int r = tid; r < KHR; r += KTPB) {
      const int i = GI(r);
      if (i >= 0 && i < nr && t.ev_zone[i]) {
        ev_consts(t, a.nu1, i, dt, a.omega_frame, evc[r][0], evc[r][1], evc[r][2], a.sg1);
        ev_consts(t, a.nu, i, dt, a.omega_frame, evc[r][0], evc[r][3], evc[r][4], a.sg2);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KPT; k++) {
    const int q = tid + k * KTPB;
    if (q >= KHR * KHC) break;
    int r = q / KHC, cc = q - r * KHC;
    int i = GI(r);
    double s = 0.0, vr = 0.0, vt = 0.0, p = 0.0, phi = 0.0;
    if (i >= 0 && i <= nr) {
      s = ls[k];
      vr = lvr[k];
      vt = lvt[k];
      if (i < nr) {
        if (a.corotating) vt = vt - a.domega * rt[RT_MED][r];
        double e = ADI ? le[k] : 0.0;
        if (a.evanescent && rt[RT_EVZ][r] != 0.0) {
          vr = damp_toward(vr, evc[r][0], evc[r][1]);
          vt = damp_toward(vt, evc[r][0], evc[r][2]);
          s = damp_toward(s, evc[r][0], rt[RT_SMED][r]);
          if (ADI) e = damp_toward(e, evc[r][0], rt[RT_EMED][r]);
        }
        if (ADI) {
          p = (a.gamma - 1.0) * e;  // Kernels.cu:196
        } else {
          const double cs = rt[RT_CS][r];
          p = s * cs * cs;
        }
        phi = cell_potential(a.P, rt[RT_MED][r], rt[RT_IMED][r], ct[0][cc], ct[1][cc], a.indirect);
      }
    }
    sS[r][cc] = s;
    sA[r][cc] = p;
    sB[r][cc] = phi;
    sC[r][cc] = vt;
    sD[r][cc] = vr;
  }
  __syncthreads();

  // ---- S1: Substep1 (Kernels.cu:13-53) on rows [1, KHR), cols [1, KHC-1) ----
  // The divisors of S1's pressure gradients, Sigma_i + Sigma_i-1 (radial)
  // and Sigma_j + Sigma_j-1 (azimuthal), reappear halved in S3; x * 2 / d is
  // x / (d / 2) (both scalings exact), so S1 divides by the halved sums and
  // keeps their reciprocal refinements (recip64) for S3 in the same thread.
  double regA[KPT], regB[KPT], hrad[KPT], hazi[KPT];
  {
#pragma unroll
    for (int k = 0; k < KPT; k++) {
      const int q = tid + k * KTPB;
      if (q >= KHR * KHC) break;
      int r = q / KHC, cc = q - r * KHC;
      int i = GI(r);
      double vri = 0.0, vti = 0.0;
      hrad[k] = hazi[k] = 0.0;
      if (r >= 1 && cc >= 1 && cc < KHC - 1 && i >= 0 && i < nr) {
        if (i > 0) {
          const recip64 hd(0.5 * (sS[r][cc] + sS[r - 1][cc]));
          hrad[k] = hd.r;
          double gradp = hd.div(sA[r][cc] - sA[r - 1][cc]) * rt[RT_IDMED][r];
          double gradphi = (sB[r][cc] - sB[r - 1][cc]) * rt[RT_IDMED][r];
          double vt2 = sC[r][cc] + sC[r][cc + 1] + sC[r - 1][cc] + sC[r - 1][cc + 1];
          vt2 = vt2 / 4.0 + a.omega_frame * rt[RT_INF][r];
          vt2 = vt2 * vt2;
          double vr_src = -gradp - gradphi;
          double vradint2 = vr_src + vt2 * rt[RT_IINF][r];
          vri = sD[r][cc] + dt * vradint2;
        }
        const double invdxtheta = rt[RT_IDXT][r];
        const recip64 hd(0.5 * (sS[r][cc] + sS[r][cc - 1]));
        hazi[k] = hd.r;
        double gradp = hd.div(sA[r][cc] - sA[r][cc - 1]) * invdxtheta;
        double gradphi = (sB[r][cc] - sB[r][cc - 1]) * invdxtheta;
        double v = sC[r][cc] - dt * (gradp + gradphi);
        v += dt * rt[RT_SUPP][r];
        vti = v;
      }
      // self-gravity kick on VradInt / VthetaInt (0 on ring 0 and outside)
      regA[k] = SGK ? vri + lkr[k] : vri;
      regB[k] = SGK ? vti + lkt[k] : vti;
    }
    if (a.opress) {  // pressure for the temperature output, own cells only
      for (int q = tid; q < KTH * KTW; q += KTPB) {
        int r = q / KTW + HALO_R, cc = q % KTW + HALO_C;
        int i = GI(r), j = j0 + cc - HALO_C;
        if (i < nr && j < ns) a.opress[(long)i * ns + j] = sA[r][cc];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KPT; k++) {
      const int q = tid + k * KTPB;
      if (q >= KHR * KHC) break;
      int r = q / KHC, cc = q - r * KHC;
      sA[r][cc] = regA[k];  // VradInt (0 on rows 0 and NRAD, as never written)
      sB[r][cc] = regB[k];  // VthetaInt
    }
  }
  __syncthreads();

  // ---- S2: viscous stress (Kernels.cu:733-771) on rows [1, KHR-1), cols [2, KHC-2) ----
  const double invdphi = t.invdphi;
  const double onethird = 1.0 / 3.0;
#pragma unroll
  for (int kq = 0; kq < KPT; kq++) {
    const int q = tid + kq * KTPB;
    if (q >= KHR * KHC) break;
    int r = q / KHC, cc = q - r * KHC;
    int i = GI(r);
    if (r < 1 || r >= KHR - 1 || cc < 2 || cc >= KHC - 2 || i < 0 || i >= nr) continue;
    const double *VrR = &sA[r][0];
    const double invRm = rt[RT_IMED][r], idsup = rt[RT_IDSUP][r];
    double drr = (sA[r + 1][cc] - VrR[cc]) * idsup;
    double dpp = (sB[r][cc + 1] - sB[r][cc]) * invdphi * invRm + 0.5 * (sA[r + 1][cc] + VrR[cc]) * invRm;
    double dv = (sA[r + 1][cc] * rt[RT_SUP][r] - VrR[cc] * rt[RT_INF][r]) * idsup * invRm;
    dv += (sB[r][cc + 1] - sB[r][cc]) * invdphi * invRm;
    const double nu = rt[RT_NU][r];
    sC[r][cc] = 2.0 * sS[r][cc] * nu * (drr - onethird * dv);
    sD[r][cc] = 2.0 * sS[r][cc] * nu * (dpp - onethird * dv);
    if (i > 0) {
      double drp = 0.5 * (rt[RT_INF][r] * (sB[r][cc] * invRm - sB[r - 1][cc] * rt[RT_IMED][r - 1]) * rt[RT_IDMED][r] +
                          (VrR[cc] - VrR[cc - 1]) * invdphi * rt[RT_IINF][r]);
      sE[r][cc] = 2.0 * 0.25 * (sS[r][cc] + sS[r - 1][cc] + sS[r][cc - 1] + sS[r - 1][cc - 1]) * nu * drp;
    }
  }
  __syncthreads();

  // ---- S3: viscous velocity update (Kernels.cu:118-150), in place ----
#pragma unroll
  for (int kq = 0; kq < KPT; kq++) {
    const int q = tid + kq * KTPB;
    if (q >= KHR * KHC) break;
    int r = q / KHC, cc = q - r * KHC;
    int i = GI(r);
    if (r < 2 || r >= KHR - 1 || cc < 3 || cc >= KHC - 3 || i < 1 || i >= nr) continue;
    // (S1 covered this cell with i > 0: both halved sums' refinements are in
    // hazi / hrad)
    if (i < nr - 1 && r < KHR - 2)
      sB[r][cc] += recip64(0.5 * (sS[r][cc] + sS[r][cc - 1]), hazi[kq])
                       .div(dt * rt[RT_IMED][r] *
                            ((rt[RT_SUP][r] * sE[r + 1][cc] - rt[RT_INF][r] * sE[r][cc]) * rt[RT_IDSUP][r] +
                             (sD[r][cc] - sD[r][cc - 1]) * invdphi + 0.5 * (sE[r][cc] + sE[r + 1][cc])));
    sA[r][cc] += recip64(0.5 * (sS[r][cc] + sS[r - 1][cc]), hrad[kq])
                     .div(dt * rt[RT_IINF][r] *
                          ((rt[RT_MED][r] * sC[r][cc] - rt[RT_MED][r - 1] * sC[r - 1][cc]) * rt[RT_IDMED][r] +
                           (sE[r][cc + 1] - sE[r][cc]) * invdphi - 0.5 * (sD[r][cc] + sD[r - 1][cc])));
  }
  __syncthreads();

  // ---- S4: artificial pressure (Kernels.cu:409-434) on rows [3, KHR-1), cols [3, KHC-4) ----
#pragma unroll
  for (int kq = 0; kq < KPT; kq++) {
    const int q = tid + kq * KTPB;
    if (q >= KHR * KHC) break;
    int r = q / KHC, cc = q - r * KHC;
    int i = GI(r);
    if (r < 3 || r >= KHR - 1 || cc < 3 || cc >= KHC - 4 || i < 0 || i >= nr) continue;
    sC[r][cc] = art_visc_q(sS[r][cc], sA[r + 1][cc] - sA[r][cc]);
    sD[r][cc] = art_visc_q(sS[r][cc], sB[r][cc + 1] - sB[r][cc]);
  }
  __syncthreads();

  // ---- S5: velocity update ("kernel", Kernels.cu:438-479), damping #2, store ----
#pragma unroll
  for (int kq = 0; kq < (KTH * KTW + KTPB - 1) / KTPB; kq++) {
    const int q = tid + kq * KTPB;
    if (q >= KTH * KTW) break;
    int r = q / KTW + HALO_R, cc = q % KTW + HALO_C;
    int i = GI(r), j = j0 + cc - HALO_C;
    if (i > nr || j >= ns) continue;
    long l = (long)i * ns + j;
    if (i == nr) {  // ghost ring: VradNew/VthetaNew are never written (0)
      a.odens[l] = sS[r][cc];
      a.ovr[l] = 0.0;
      a.ovt[l] = 0.0;
      continue;
    }
    double vrnew = 0.0;
    if (i > 0) {
      double d = sS[r][cc] + sS[r - 1][cc];
      double densint = sC[r][cc] - sC[r - 1][cc];
      vrnew = sA[r][cc] - dt * 2.0 / d * densint * rt[RT_IDMED][r];
    }
    const double invdxtheta = rt[RT_IDXT][r];
    double dens2 = sS[r][cc] + sS[r][cc - 1];
    double tempint = (sD[r][cc] - sD[r][cc - 1]);
    double vtnew = sB[r][cc] - dt * 2.0 / dens2 * tempint * invdxtheta;
    double s = sS[r][cc];
    if constexpr (ADI) {
      // EnergyInt ("kernel", Kernels.cu:474-477) from the energy after the
      // first damping pass; the second pass damps Energy, which Substep3
      // then overwrites, so it is not formed
      double e = a.energy[l];
      if (a.evanescent && rt[RT_EVZ][r] != 0.0) e = damp_toward(e, evc[r][0], rt[RT_EMED][r]);
      a.oeint[l] = e - dt * sC[r][cc] * (sA[r + 1][cc] - sA[r][cc]) * rt[RT_IDSUP][r] -
                   dt * sD[r][cc] * (sB[r][cc + 1] - sB[r][cc]) * invdxtheta;
    }
    if (a.evanescent && rt[RT_EVZ][r] != 0.0) {
      vrnew = damp_toward(vrnew, evc[r][0], evc[r][3]);
      vtnew = damp_toward(vtnew, evc[r][0], evc[r][4]);
      s = damp_toward(s, evc[r][0], rt[RT_SMED][r]);
    }
    a.odens[l] = s;
    a.ovr[l] = vrnew;
    a.ovt[l] = vtnew;
  }
#undef GI
#undef GJ
}

// ----------------------------------------------------------- k_src3 ----
// Substep3 of an adiabatic deck (SourceEuler.cu:287-291): the viscous
// terms of the velocities after the second damping pass (ViscousTermsKernel,
// Kernels.cu:733-771), Qplus (Substep3Kernel, Kernels.cu:56-77; ring 0 by
// the power-law extrapolation from rings 1 and 2, Kernels.cu:88-95) and
// the implicit energy update with optional relaxation cooling
// (Substep3Kernel2, Kernels.cu:97-113).  One thread per cell; the
// neighbours are read through the caches.
struct src3_args {
  const double *dens, *vr, *vt, *eint;
  double *energy;
  const double *nu;
  fg_dev_tables t;
  double dt, gamma, lr01, lr12;
  int nr, ns, cooling;
};

// divV of cell (i, j) and, for i >= 1, its Qplus
__device__ __forceinline__ void src3_cell(const src3_args &a, int i, int j, double &qplus, double &divv) {
  const fg_dev_tables &t = a.t;
  const int ns = a.ns;
  const long l = (long)i * ns + j;
  const long JP = (long)i * ns + (j + 1 == ns ? 0 : j + 1), JM = (long)i * ns + (j == 0 ? ns - 1 : j - 1);
  const double invdphi = t.invdphi;
  const double onethird = 1.0 / 3.0;
  const double vrl = a.vr[l], vrn = a.vr[l + ns], vtl = a.vt[l], vtp = a.vt[JP];
  const double drr = (vrn - vrl) * t.invdiffRsup[i];
  const double dpp = (vtp - vtl) * invdphi * t.invRmed[i] + 0.5 * (vrn + vrl) * t.invRmed[i];
  double dv = (vrn * t.Rsup[i] - vrl * t.Rinf[i]) * t.invdiffRsup[i] * t.invRmed[i];
  dv += (vtp - vtl) * invdphi * t.invRmed[i];
  divv = dv;
  qplus = 0.0;
  if (i == 0) return;
  const double d = a.dens[l], nu = a.nu[i];
  if (nu == 0.0) return;
  const double TRR = 2.0 * d * nu * (drr - onethird * dv);
  const double TPP = 2.0 * d * nu * (dpp - onethird * dv);
  const double drp = 0.5 * (t.Rinf[i] * (vtl * t.invRmed[i] - a.vt[l - ns] * t.invRmed[i - 1]) * t.invdiffRmed[i] +
                            (vrl - a.vr[JM]) * invdphi * t.invRinf[i]);
  const double TRP = 2.0 * 0.25 * (d + a.dens[l - ns] + a.dens[JM] + a.dens[JM - ns]) * nu * drp;
  double q = 0.5 / nu / d * (TRR * TRR + TRP * TRP + TPP * TPP);
  q += (2.0 / 9.0) * nu * d * dv * dv;
  qplus = q;
}

__global__ __launch_bounds__(256) void k_src3(src3_args a) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= a.nr || j >= a.ns) return;
  const fg_dev_tables &t = a.t;
  const long l = (long)i * a.ns + j;
  double qplus, divv;
  src3_cell(a, i, j, qplus, divv);
  if (i == 0) {
    if (a.nu[a.nr - 1] != 0.0) {
      double q1, q2, unused;
      src3_cell(a, 1, j, q1, unused);
      src3_cell(a, 2, j, q2, unused);
      qplus = q1 * exp(log(q1 / q2) * a.lr01 / a.lr12);
    } else {
      qplus = 0.0;
    }
  }
  double num, den;
  if (!a.cooling) {
    num = a.dt * qplus + a.eint[l];
    den = 1.0 + (a.gamma - 1.0) * a.dt * divv;
  } else {
    const double d = a.dens[l];
    num = t.EnergyMed[i] * a.dt * d / t.SigmaMed[i] + t.CoolingTimeMed[i] * a.eint[l] +
          a.dt * t.CoolingTimeMed[i] * (qplus - t.QplusMed[i] * d / t.SigmaMed[i]);
    den = a.dt + t.CoolingTimeMed[i] + (a.gamma - 1.0) * a.dt * t.CoolingTimeMed[i] * divv;
  }
  a.energy[l] = num / den;
}

// ----------------------------------------------------------- k_trad ----
// Radial sweep (OneWindRad, TransportEuler.cu:60-79) marching over rings
// of a band; one thread owns one sector column.  For each ring m the
// momenta (LRMomentaKernel), Work = Q / (Sigma + 1e-20), the slope and
// the face value at face m-1 are formed once and kept in registers; the
// face flux F(k) = dt*dtheta*Rinf[k]*Q*(k)*Sigma*(k)*vrad(k) serves both
// neighbouring rings (Rsup[k-1] == Rinf[k]).
// FP32 bulk mode storage.  The fields passed between the fused kernels
// (after k_src, k_trad and k_taz) are stored in FP32 as offsets from FP64
// backgrounds that writer and reader form alike: Sigma from sbar (ring
// means of the state the substep starts from), vtheta from vref (the ring
// means of vtheta at that state, Vmoy), the angular momenta from the
// cell's own stored Sigma times (vref + R Omega) R (fp32_bg_angmom; the
// writer takes Sigma as rounded for the reader); vrad and the radial
// momenta are stored as they are.  The offsets are small against the
// fields (a few 1e-3 of vtheta, the non-axisymmetric part of Sigma), so
// their FP32 rounding stays far below the FP32 arithmetic's own error; the
// state between substeps stays FP64.
// Near the planets, where circumplanetary material stays for many orbits
// and the rounding of the stored Sigma would add up there, the rings
// [lob0, lob1) store Sigma as a float pair: the high part where every ring
// stores its float (index l < S of an array of S doubles) and the rounding
// residual at S + l, in the array's otherwise unused upper half; a reader
// adds the residual only on those rings.  (Pairs for the velocities and
// momenta as well change nothing measurable.)  The upper-half slots of the
// ghost ring NRAD are never written and stay zero: loads that must be
// issued on every iteration take them for rings outside the band.
__device__ __forceinline__ bool in_lob(int i, int b0, int b1) { return i >= b0 && i < b1; }
__device__ __forceinline__ double rec32(double bg, float hi, float lo) { return bg + ((double)hi + (double)lo); }
__device__ __forceinline__ double fp32_bg_angmom(double sbar, double vref, double R, double omega) {
  return sbar * (vref + R * omega) * R;
}

#define TRAD_TPB 64
#define NQ 5  // RadMomP, RadMomM, ThetaMomP, ThetaMomM, Sigma
#define NQ_MAX 6  // adiabatic: RadMomP, RadMomM, ThetaMomP, ThetaMomM, Energy, Sigma

struct trad_args {
  const double *dens, *vr, *vt;
  const double *energy;  // adiabatic decks (NQT = 6)
  double *out[NQ_MAX];
  // FP32 bulk mode: the same fields stored in FP32 (fp32_store_bg)
  const float *fdens, *fvr, *fvt;
  float *fout[NQ];
  const double *sbar, *vref;  // per-ring backgrounds of Sigma and vtheta
  double *vmed;  // ring means of vtheta for the azimuthal pass
  fg_dev_tables t;
  double dt, omega_frame;
  int nr, ns, nstrips;
  int band;        // rings per column strip
  int lob0, lob1;  // FP32 storage: rings stored as float pairs
};

// ComputeAverageThetaVelocitiesKernel (Kernels.cu:1070-1081) for one ring
// by one wave: the lanes stage the ring through LDS 512 sectors at a time
// (the next 512 loaded into registers meanwhile) and lane 0 adds them in j
// order with its LDS reads issued ahead of the dependent adds.  One wave:
// the barrier only orders the LDS accesses, and lane 0 has finished a
// round's sum before the next round's stores in program order.
#define RMW_CH 8  // sectors per lane per round
__device__ void ring_mean_wave(const double *v, int ns, double *out) {
  __shared__ double buf[64 * RMW_CH];
  const int lane = threadIdx.x;
  const int nch = (ns + 64 * RMW_CH - 1) / (64 * RMW_CH);
  double nx[RMW_CH];
#pragma unroll
  for (int u = 0; u < RMW_CH; u++) nx[u] = (lane + 64 * u < ns) ? v[lane + 64 * u] : 0.0;
  double sum = 0.0;
  for (int c = 0; c < nch; c++) {
#pragma unroll
    for (int u = 0; u < RMW_CH; u++) buf[lane + 64 * u] = nx[u];
    const int e0 = (c + 1) * 64 * RMW_CH + lane;
#pragma unroll
    for (int u = 0; u < RMW_CH; u++) nx[u] = (e0 + 64 * u < ns) ? v[e0 + 64 * u] : 0.0;
    __syncthreads();
    if (lane == 0) sum = seq_sum_lds(buf, min(64 * RMW_CH, ns - c * 64 * RMW_CH), sum);
    __syncthreads();
  }
  if (lane == 0) *out = sum / (double)ns;
}

// x of lane `lane` (uniform), as scalars
__device__ __forceinline__ double lane_get(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Ring tables of the rings a column strip marches over, one ring per lane
// (lane r holds ring r0 + r, clamped into the nr + 1 entries), read back
// with v_readlane.  Reading the tables from memory inside the ring loop
// would make each iteration wait for that load and, as a wave's vector
// memory operations complete in order, for the next ring's field loads
// issued before it, leaving nothing in flight across the arithmetic.
struct ring_lanes {
  int r0;
  double Rmed, idm, Rinf, isurf;
  __device__ __forceinline__ void load(const fg_dev_tables &t, int base, int nr) {
    r0 = base;
    const int r = min(max(base + (int)threadIdx.x, 0), nr);
    Rmed = t.Rmed[r];
    idm = t.invdiffRmed[r];
    Rinf = t.Rinf[r];
    isurf = t.invSurf[r];
    // consume the loads here, so that no wait for them lands in the loop
    asm volatile("" ::"v"(Rmed), "v"(idm), "v"(Rinf), "v"(isurf));
  }
  // ring m - 2 .. m + 61 can be read: reload when ring m would leave that
  __device__ __forceinline__ int at(const fg_dev_tables &t, int m, int nr) {
    if (m - r0 >= 64) load(t, m - 2, nr);
    return m - r0;
  }
};

// The FP32 storage backgrounds of the rings a column strip marches over,
// one ring per lane like ring_lanes (same first ring r0, reloaded with it)
struct bg_lanes {
  int r0;
  double sb, vb;
  __device__ __forceinline__ void load(const double *sbar, const double *vref, int base, int nr) {
    r0 = base;
    const int r = min(max(base + (int)threadIdx.x, 0), nr);
    sb = sbar[r];
    vb = vref[r];
    asm volatile("" ::"v"(sb), "v"(vb));
  }
};

// FP32 mode's VMed (ComputeAverageThetaVelocitiesKernel, Kernels.cu:
// 1070-1081) of one ring by one wave: vtheta (background + stored offset)
// summed per lane over sectors lane, lane + 64, ... (eight loads in flight),
// then a shuffle tree.  The FP32 mode does not replay the reference's
// one-thread order; the sum is no less accurate.
__device__ void ring_mean_wave32(const float *v, double vb, int ns, double *out) {
  const int lane = threadIdx.x;
  double sum = 0.0;
  for (int j0 = lane; j0 < ns; j0 += 64 * 8) {
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = (j0 + 64 * u < ns) ? v[j0 + 64 * u] : 0.0f;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + 64 * u < ns) sum += vb + (double)x[u];
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
  if (lane == 0) *out = sum / (double)ns;
}

// NQT transported quantities: 5, or 6 with the energy of an adiabatic deck
// (TransportEuler.cu:60-79 sweeps it between the momenta and Sigma)
template <int NQT>
__global__ __launch_bounds__(TRAD_TPB) void k_trad(trad_args a) {
  constexpr bool EN = NQT == 6;
  const int nr = a.nr, ns = a.ns;
  // blocks [0, NRAD): one ring mean each, issued first so that these
  // latency-bound sums overlap the sweep blocks
  if ((int)blockIdx.x < nr) {
    ring_mean_wave(a.vt + (long)blockIdx.x * ns, ns, a.vmed + blockIdx.x);
    return;
  }
  const int bid = blockIdx.x - nr;
  // lanes past the last sector march on column ns - 1 (every lane holds
  // its ring table entry) and store the same values as the lane that owns
  // that column
  const int j = min((bid % a.nstrips) * TRAD_TPB + (int)threadIdx.x, ns - 1);
  const int jp = (j + 1) % ns;
  const int b0 = (bid / a.nstrips) * a.band;
  const int b1 = min(b0 + a.band, nr);
  const fg_dev_tables &t = a.t;
  const double dt = a.dt, OF = a.omega_frame;
  const double cflux = dt * t.dtheta;
  // windows over the last three rings: ring x (its Q, W, Sigma, vrad and
  // slope) and face x (its flux) live in slot x % 3, so a step writes its
  // new ring over the one three back and nothing is shifted; the step for
  // ring m is instantiated per m % 3 to keep the slots static
  double Q[NQT][3] = {}, W[NQT][3] = {}, dq[NQT][3] = {}, F[NQT][3] = {};
  double sg[3] = {}, dqs[3] = {}, vr[3] = {};
  const int m0 = max(b0 - 2, 0);
  ring_lanes rt;
  rt.load(t, m0 - 2, nr);
  // ring m's inputs are loaded one iteration ahead (ring min(m + 1, nr - 1):
  // the loads are unconditional, so that every iteration of a loop issues
  // the same memory operations and the waits for them can count past the
  // stores); vrad of ring m + 1 is carried over as the next ring's own
  double n_s, n_v0, n_v1, n_vtp, n_vt, n_e = 0.0;
  {
    const long l = (long)m0 * ns + j;
    n_s = a.dens[l];
    n_v0 = a.vr[l];
    n_v1 = a.vr[l + ns];
    n_vtp = a.vt[(long)m0 * ns + jp];
    n_vt = a.vt[l];
    if constexpr (EN) n_e = a.energy[l];
    asm volatile("" ::"v"(n_s), "v"(n_v0), "v"(n_v1), "v"(n_vtp), "v"(n_vt), "v"(n_e));
  }
  // inner: the caller guarantees 2 <= m < nr (ring m inside the disk, slope
  // and face of ring m - 1 interior), which drops the edge cases
  auto ring = [&](int m, auto store, auto phase, auto inner) {
    constexpr int C = decltype(phase)::value, B1 = (C + 2) % 3, B2 = (C + 1) % 3;  // rings m, m-1, m-2
    constexpr bool IN = decltype(inner)::value;
    const int r = rt.at(t, m, nr);
    const double c_s = n_s, c_v0 = n_v0, c_v1 = n_v1, c_vtp = n_vtp, c_vt = n_vt, c_e = n_e;
    {
      const long l = (long)min(m + 1, nr - 1) * ns;
      n_s = a.dens[l + j];
      n_v0 = c_v1;
      n_v1 = a.vr[l + ns + j];
      n_vtp = a.vt[l + jp];
      n_vt = a.vt[l + j];
      if constexpr (EN) n_e = a.energy[l + j];
    }
    // ring tables: Rmed[m], Rmed[m-1], Rmed[m-2] at lanes r, r-1, r-2
    const double Rm = lane_get(rt.Rmed, r);
    if (IN || m < nr) {
      const double s = c_s;
      const double v0 = c_v0, v1 = c_v1;
      Q[0][C] = s * v1;
      Q[1][C] = s * v0;
      Q[2][C] = s * (c_vtp + Rm * OF) * Rm;
      Q[3][C] = s * (c_vt + Rm * OF) * Rm;
      if constexpr (EN) Q[4][C] = c_e;
      Q[NQT - 1][C] = s;
      sg[C] = s;
      vr[C] = v0;
      const recip64 rs(s + 1e-20);  // Work = Q / (Sigma + 1e-20), one denominator
      for (int q = 0; q < NQT; q++) W[q][C] = rs.div(Q[q][C]);
    } else {
      for (int q = 0; q < NQT; q++) Q[q][C] = W[q][C] = 0.0;
      sg[C] = 0.0;
      vr[C] = 0.0;  // vrad of the ghost ring
    }
    // slope of ring k = m-1 (zero on rings 0 and NRAD-1)
    const int k = m - 1;
    if (!decltype(store)::value && k < 0) return;  // storing steps have k >= 1
    const bool flat = !IN && (k == 0 || k >= nr - 1);
    const double idm0 = lane_get(rt.idm, r - 1), idm1 = lane_get(rt.idm, r);
    for (int q = 0; q < NQT; q++) dq[q][B1] = flat ? 0.0 : vl_rad_slope(W[q][B2], W[q][B1], W[q][C], idm0, idm1);
    dqs[B1] = flat ? 0.0 : vl_rad_slope(sg[B2], sg[B1], sg[C], idm0, idm1);
    // face k: upwind values (StarRadKernel2); faces 0 and NRAD are zero
    double DS = 0.0;
    double QS[NQT];
    if (IN || (k > 0 && k < nr)) {
      const double vk = vr[B1];
      const double Rkm = lane_get(rt.Rmed, r - 2), Rk = lane_get(rt.Rmed, r - 1);
      DS = vl_rad_star(vk, dt, sg[B2], sg[B1], dqs[B2], dqs[B1], Rkm, Rk, Rm, k);
      for (int q = 0; q < NQT; q++)
        QS[q] = vl_rad_star(vk, dt, W[q][B2], W[q][B1], dq[q][B2], dq[q][B1], Rkm, Rk, Rm, k);
    } else {
      for (int q = 0; q < NQT; q++) QS[q] = 0.0;
    }
    const double vface = (IN || k < nr) ? vr[B1] : 0.0;
    const double Rinf = lane_get(rt.Rinf, r - 1);
    for (int q = 0; q < NQT; q++) F[q][B1] = cflux * Rinf * QS[q] * DS * vface;
    // ring k-1 is complete: Q += (F(k-1) - F(k)) / Surf
    if (decltype(store)::value) {
      const int o = k - 1;
      const long lo = (long)o * ns + j;
      const double isurf = lane_get(rt.isurf, r - 2);
      for (int q = 0; q < NQT; q++) {
        double varq = F[q][B2];
        varq -= F[q][B1];
        a.out[q][lo] = Q[q][B2] + varq * isurf;
      }
    }
  };
  // a single step in the rotating layout (newest ring in slot 2): shift
  // every window by one, then the step of phase 2.  Three consecutive
  // steps of phases 0, 1, 2 leave the windows in that same layout.
  auto shift = [&]() {
    for (int q = 0; q < NQT; q++) {
      Q[q][0] = Q[q][1];
      Q[q][1] = Q[q][2];
      W[q][0] = W[q][1];
      W[q][1] = W[q][2];
      dq[q][0] = dq[q][1];
      F[q][0] = F[q][1];
    }
    sg[0] = sg[1];
    sg[1] = sg[2];
    vr[0] = vr[1];
    vr[1] = vr[2];
    dqs[0] = dqs[1];
  };
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;
  using P2 = std::integral_constant<int, 2>;
  // rings m0 .. b0 + 1 fill the window; from m = b0 + 2 on ring m - 2 of
  // the band is complete after every step
  using T = std::true_type;
  using F0 = std::false_type;
  int m = m0;
  for (; m < b0 + 2; m++) {
    shift();
    ring(m, F0(), P2(), F0());
  }
  // the storing loop starts with nothing in flight from the filling one;
  // its steps have m >= b0 + 2 >= 2, so those below nr are inner
  asm volatile("" ::"v"(n_s), "v"(n_v1), "v"(n_vtp), "v"(n_vt), "v"(n_e));
  const int mi = min(b1 + 1, nr - 1);
  for (; m + 2 <= mi; m += 3) {
    ring(m, T(), P0(), T());
    ring(m + 1, T(), P1(), T());
    ring(m + 2, T(), P2(), T());
  }
  asm volatile("" ::"v"(n_s), "v"(n_v1), "v"(n_vtp), "v"(n_vt), "v"(n_e));
  for (; m <= b1 + 1; m++) {
    shift();
    ring(m, T(), P2(), F0());
  }
}

// ------------------------------------------------------------ k_taz ----
// OneWindTheta (TransportEuler.cu:177-194) + the vtheta half of
// ComputeVelocities (Kernels.cu:1206-1207) as wave chunks (k_tazw below),
// FP64 and FP32 bulk mode.  VMed comes from the ring-mean waves of k_trad.

struct taz_args {
  const double *vt;     // vtheta after k_src
  const double *vmed;   // its ring means
  const double *q[NQ_MAX];  // quantities after the radial sweep (Sigma last)
  double *xd, *xvt, *xrmp, *xrmm, *xe;
  // FP32 bulk mode: the same fields stored in FP32 (fp32_store_bg)
  const float *fvt, *fq[NQ];
  float *fxd, *fxvt, *fxrmp, *fxrmm;
  const double *sbar, *vref;
  int *nshift;
  fg_dev_tables t;
  double dt, omega_frame;
  int nr, ns, fast;
  int lob0, lob1;  // FP32 storage: rings stored as float pairs
};


// ------------------------------------------------------- k_taz, waves ----
// No workgroup barrier: one 64-lane wave owns a
// contiguous chunk of 64*CL cells of a ring, CL consecutive cells per lane
// in registers.  A cell's neighbours are in the lane's own registers, or for
// a lane's first and last cells in the adjacent lane, fetched by DPP wave
// shifts.  The chunk carries TAZW_HL halo cells on the left and TAZW_HR on
// the right: each sweep's result at a cell depends on the cells up to two
// away, the vtheta recovery on the cell to the left, and the two lanes at
// the chunk's ends see zeros from the shifts, so the outermost 4 (right)
// and 5 (left) results of a chunk are discarded.  Chunks of one ring
// overlap by those halos; each cell is output by exactly one chunk.
#define TAZW_HL 5
#define TAZW_HR 4
#define TAZW_P32 (64 * 4 - TAZW_HL - TAZW_HR)  // cells output per batched FP32 chunk (CL = 4)
// single-disk FP32 chunks: three cells per lane in their own kernel held to
// 96 VGPRs, five waves per SIMD (seven dwords go to scratch)
#define TAZW_CL32 3
#define TAZW_P32S (64 * TAZW_CL32 - TAZW_HL - TAZW_HR)
// FP64 chunks: three cells per lane (the five- and six-quantity sweeps fit
// in 142 and 164 VGPRs, three waves per SIMD; four or five cells per lane
// take 174-208 VGPRs and leave two)
#define TAZW_CL64 3
#define TAZW_P64 (64 * TAZW_CL64 - TAZW_HL - TAZW_HR)  // cells output per FP64 chunk

// x of lane l-1 (lane 0 reads 0): DPP wave_shr:1 on each dword
__device__ __forceinline__ double lane_prev(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// x of lane l+1 (lane 63 reads 0): DPP wave_shl:1 on each dword
__device__ __forceinline__ double lane_next(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float lane_prev(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float lane_next(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}

// QuantitiesAdvection (TransportEuler.cu:229-245) on a lane's CL cells with
// velocity V (one value per cell, or the ring's uniform residual when
// UNIFORM); the reference's operations in its order
template <int CL, bool UNIFORM, int NQT>
__device__ __forceinline__ void wave_sweep(double (&q)[NQT][CL], const double (&Vc)[CL], double Vu,
                                           const fg_dev_tables &t, int i, double dt) {
  const double invdxtheta = t.invdxtheta[i], dxtheta = t.dxtheta[i];
  const double dxrad = (t.Rsup[i] - t.Rinf[i]) * dt;
  const double invsurf = t.invSurf[i];
#define V(c) (UNIFORM ? Vu : Vc[c])
  double dens[CL], DS[CL];
  // DensStar from Sigma (ComputeStarTheta(Dens), TransportEuler.cu:232)
#pragma unroll
  for (int c = 0; c < CL; c++) dens[c] = q[NQT - 1][c];
  {
    const double Wm = lane_prev(dens[CL - 1]), Wp = lane_next(dens[0]);
    double D[CL];
#pragma unroll
    for (int c = 0; c < CL; c++)
      D[c] = vl_theta_slope(c ? dens[c - 1] : Wm, dens[c], c < CL - 1 ? dens[c + 1] : Wp, invdxtheta);
    const double Dm = lane_prev(D[CL - 1]);
#pragma unroll
    for (int c = 0; c < CL; c++)
      DS[c] = vl_theta_star(V(c), dt, c ? dens[c - 1] : Wm, dens[c], c ? D[c - 1] : Dm, D[c], dxtheta);
  }
  // VanLeerTheta of each quantity, Sigma last (TransportEuler.cu:234-243);
  // the Work divisions of a cell share its denominator Sigma + 1e-20
  double rd[CL];
#pragma unroll
  for (int c = 0; c < CL; c++) rd[c] = recip64(dens[c] + 1e-20).r;
#pragma unroll
  for (int qq = 0; qq < NQT; qq++) {
    // one quantity at a time: keeps the live registers to this quantity's
    // stencil (the scheduler would otherwise interleave all five)
    __builtin_amdgcn_sched_barrier(0);
    double W[CL], D[CL], F[CL];
#pragma unroll
    for (int c = 0; c < CL; c++) W[c] = recip64(dens[c] + 1e-20, rd[c]).div(q[qq][c]);
    const double Wm = lane_prev(W[CL - 1]), Wp = lane_next(W[0]);
#pragma unroll
    for (int c = 0; c < CL; c++)
      D[c] = vl_theta_slope(c ? W[c - 1] : Wm, W[c], c < CL - 1 ? W[c + 1] : Wp, invdxtheta);
    const double Dm = lane_prev(D[CL - 1]);
#pragma unroll
    for (int c = 0; c < CL; c++) {
      const double QS = vl_theta_star(V(c), dt, c ? W[c - 1] : Wm, W[c], c ? D[c - 1] : Dm, D[c], dxtheta);
      F[c] = dxrad * QS * DS[c] * V(c);
    }
    const double Fp = lane_next(F[0]);
#pragma unroll
    for (int c = 0; c < CL; c++) {
      double varq = F[c];
      varq -= (c < CL - 1 ? F[c + 1] : Fp);
      q[qq][c] += varq * invsurf;
    }
  }
#undef V
}

// The FP32 bulk-mode sweep on a lane's CL cells: W, slopes, upwind values
// and fluxes in FP32 with hardware reciprocals, each increment added to the
// FP64 quantity
template <int CL, bool UNIFORM>
__device__ __forceinline__ void wave_sweep32(double (&q)[NQ][CL], const float (&Vc)[CL], float Vu,
                                             const fg_dev_tables &t, int i, float dt) {
  const float invdxtheta = (float)t.invdxtheta[i], dxtheta = (float)t.dxtheta[i];
  const float dxrad = (float)((t.Rsup[i] - t.Rinf[i]) * (double)dt);
  const float invsurf = (float)t.invSurf[i];
#define V(c) (UNIFORM ? Vu : Vc[c])
  float rs[CL], DS[CL];
  {
    float d[CL], D[CL];
#pragma unroll
    for (int c = 0; c < CL; c++) {
      d[c] = (float)q[NQ - 1][c];
      rs[c] = rcp_f(d[c] + 1e-20f);
    }
    const float Wm = lane_prev(d[CL - 1]), Wp = lane_next(d[0]);
#pragma unroll
    for (int c = 0; c < CL; c++)
      D[c] = vl_theta_slope_f(c ? d[c - 1] : Wm, d[c], c < CL - 1 ? d[c + 1] : Wp, invdxtheta);
    const float Dm = lane_prev(D[CL - 1]);
#pragma unroll
    for (int c = 0; c < CL; c++)
      DS[c] = vl_theta_star_f(V(c), dt, c ? d[c - 1] : Wm, d[c], c ? D[c - 1] : Dm, D[c], dxtheta) * V(c) * dxrad;
  }
#pragma unroll
  for (int qq = 0; qq < NQ; qq++) {
    __builtin_amdgcn_sched_barrier(0);
    float W[CL], D[CL], F[CL];
#pragma unroll
    for (int c = 0; c < CL; c++) W[c] = (float)q[qq][c] * rs[c];
    const float Wm = lane_prev(W[CL - 1]), Wp = lane_next(W[0]);
#pragma unroll
    for (int c = 0; c < CL; c++)
      D[c] = vl_theta_slope_f(c ? W[c - 1] : Wm, W[c], c < CL - 1 ? W[c + 1] : Wp, invdxtheta);
    const float Dm = lane_prev(D[CL - 1]);
#pragma unroll
    for (int c = 0; c < CL; c++)
      F[c] = vl_theta_star_f(V(c), dt, c ? W[c - 1] : Wm, W[c], c ? D[c - 1] : Dm, D[c], dxtheta) * DS[c];
    const float Fp = lane_next(F[0]);
#pragma unroll
    for (int c = 0; c < CL; c++) q[qq][c] += (double)((F[c] - (c < CL - 1 ? F[c + 1] : Fp)) * invsurf);
  }
#undef V
}

// chunk `chunk` of ring i, one wave; F32 selects the FP32 bulk-mode sweeps;
// NQT = 6 adds the energy of an adiabatic deck (FP64 only)
// LDS of one chunk: the load stage, then the output stages (row 0 reused)
template <int CL, int NQT>
struct tazw_lds {
  double ostage[NQT == 6 ? 5 : 4][64 * CL];
};
// PAIRS: a ring of the FP32 mode stored as float pairs (its own instance,
// so that the other rings run without the residuals' loads and branches)
template <int CL, bool F32, int NQT = 5, bool PAIRS = false>
__device__ __forceinline__ void tazw_body(const taz_args &a, int i, int chunk, tazw_lds<CL, NQT> &lds) {
  constexpr int N = 64 * CL, P = N - TAZW_HL - TAZW_HR;
  static_assert(NQT == 5 || !F32, "the FP32 sweeps carry five quantities");
  static_assert(F32 || !PAIRS, "float pairs are an FP32 storage form");
  auto &ostage = lds.ostage;
  double *stage = ostage[0];
  const int ns = a.ns, l = threadIdx.x;
  const int g0 = chunk * P - TAZW_HL;  // unwrapped index of the chunk's first cell
  const fg_dev_tables &t = a.t;
  const long base = (long)i * ns;
  // loads: coalesced global reads into the chunk's LDS stage, then CL
  // consecutive cells per lane (a one-wave workgroup: the barriers only
  // order the LDS accesses of the wave)
  double q[NQT][CL];
  using vel_t = typename std::conditional<F32, float, double>::type;
  vel_t vres[CL];
  const double VMed = a.vmed[i];
  double vt0[CL];
  // ring index of this lane's staged cell m: one modulo, then steps of 64
  int jl[CL];
  {
    int j = ((g0 + l) % ns + ns) % ns;
#pragma unroll
    for (int m = 0; m < CL; m++) {
      jl[m] = j;
      j += 64;
      while (j >= ns) j -= ns;
    }
  }
  // every field's loads are issued before the first is staged, so the six
  // fields cost one memory latency, not six.  FP32 mode: the fields are
  // FP32 offsets from the ring's backgrounds (fp32_store_bg)
  double raw[NQT + 1][CL];
  double sb = 0.0, vb = 0.0, lb = 0.0;
  constexpr bool lob = PAIRS;  // FP32 storage: this ring is stored as float pairs
  const long S = (long)(a.nr + 1) * ns;
  if constexpr (F32) {
    float fr[NQT + 1][CL], fsl[CL];
#pragma unroll
    for (int qq = 0; qq <= NQT; qq++) {
      const float *src = qq < NQT ? a.fq[qq] : a.fvt;
#pragma unroll
      for (int m = 0; m < CL; m++) fr[qq][m] = src[base + jl[m]];
    }
#pragma unroll
    for (int m = 0; m < CL; m++) fsl[m] = lob ? a.fq[4][S + base + jl[m]] : 0.f;
    sb = a.sbar[i];
    vb = a.vref[i];
    const double Rm = t.Rmed[i];
#pragma unroll
    for (int m = 0; m < CL; m++) {
      const double s = rec32(sb, fr[4][m], fsl[m]);
      lb = fp32_bg_angmom(s, vb, Rm, a.omega_frame);  // the cell's angular momentum background
      raw[0][m] = (double)fr[0][m];
      raw[1][m] = (double)fr[1][m];
      raw[2][m] = lb + (double)fr[2][m];
      raw[3][m] = lb + (double)fr[3][m];
      raw[4][m] = s;
      raw[5][m] = vb + (double)fr[5][m];
    }
  } else {
#pragma unroll
    for (int qq = 0; qq <= NQT; qq++) {
      const double *src = qq < NQT ? a.q[qq] : a.vt;
#pragma unroll
      for (int m = 0; m < CL; m++) raw[qq][m] = src[base + jl[m]];
    }
  }
#pragma unroll
  for (int qq = 0; qq <= NQT; qq++) {
#pragma unroll
    for (int m = 0; m < CL; m++) stage[l + 64 * m] = raw[qq][m];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CL; c++) {
      const double v = stage[l * CL + c];
      if (qq < NQT)
        q[qq][c] = v;
      else
        vt0[c] = v - VMed;  // ComputeResidualsKernel
    }
    __syncthreads();
  }
  // ComputeConstantResidualKernel (Kernels.cu:1094-1126)
  const double invdt = 1.0 / a.dt;
  const double dpinvns = 2.0 * FG_PI / (double)ns;
  const double Ntilde = VMed * t.invRmed[i] * a.dt * (double)ns / 2.0 / FG_PI;
  const double Nround = floor(Ntilde + 0.5);
  const int nsh = (int)(long)Nround;
  double vuni = (Ntilde - Nround) * t.Rmed[i] * invdt * dpinvns;
#pragma unroll
  for (int c = 0; c < CL; c++) vres[c] = (vel_t)(a.fast ? vt0[c] : vt0[c] + vuni);
  if (!a.fast) vuni = 0.0;
  if (chunk == 0 && l == 0) a.nshift[i] = nsh;
  if constexpr (F32 && NQT == 5) {
    wave_sweep32<CL, false>(q, vres, 0.0f, t, i, (float)a.dt);
    wave_sweep32<CL, true>(q, vres, (float)vuni, t, i, (float)a.dt);
  } else {
    wave_sweep<CL, false, NQT>(q, vres, 0.0, t, i, a.dt);
    wave_sweep<CL, true, NQT>(q, vres, vuni, t, i, a.dt);
  }
  // vtheta from the angular momenta (Kernels.cu:1206-1207): cells j - 1
  // and j of the unshifted ring
  const double Rm = t.Rmed[i];
  const recip64 rRm(Rm);
  double vt[CL];
  {
    const double Wm = lane_prev(q[2][CL - 1]), Sm = lane_prev(q[NQT - 1][CL - 1]);
#pragma unroll
    for (int c = 0; c < CL; c++) {
      const double s = q[NQT - 1][c];
      vt[c] = rRm.div(((c ? q[2][c - 1] : Wm) + q[3][c]) / (s + (c ? q[NQT - 1][c - 1] : Sm) + 1e-15)) -
              Rm * a.omega_frame;
    }
  }
  // stores with AdvectSHIFT (Kernels.cu:1173-1190) folded in: cell j goes
  // to j + Nshift (mod NSEC); the chunk outputs its cells [HL, HL + P)
  // that lie in [0, NSEC)
  int sh = nsh % ns;
  if (sh < 0) sh += ns;
  // all four outputs staged at once, then their coalesced stores back to
  // back (one LDS round trip instead of four)
#pragma unroll
  for (int c = 0; c < CL; c++) {
    ostage[0][l * CL + c] = q[NQT - 1][c];
    ostage[1][l * CL + c] = vt[c];
    ostage[2][l * CL + c] = q[0][c];
    ostage[3][l * CL + c] = q[1][c];
    if constexpr (NQT == 6) ostage[4][l * CL + c] = q[4][c];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < CL; m++) {
    const int k = l + 64 * m;
    const int g = g0 + k;
    if (k >= TAZW_HL && k < TAZW_HL + P && g < ns) {
      int jo = g + sh;
      if (jo >= ns) jo -= ns;
      const long lo = base + jo;
      if constexpr (F32) {
        const double xd = ostage[0][k] - sb;
        const float hd = (float)xd;
        a.fxd[lo] = hd;
        a.fxvt[lo] = (float)(ostage[1][k] - vb);
        a.fxrmp[lo] = (float)ostage[2][k];
        a.fxrmm[lo] = (float)ostage[3][k];
        if constexpr (lob) a.fxd[S + lo] = (float)(xd - (double)hd);
      } else {
        a.xd[lo] = ostage[0][k];
        a.xvt[lo] = ostage[1][k];
        a.xrmp[lo] = ostage[2][k];
        a.xrmm[lo] = ostage[3][k];
        if constexpr (NQT == 6) a.xe[lo] = ostage[4][k];
      }
    }
  }
}

// blockIdx.x = chunk, blockIdx.y = ring
template <int CL, bool F32, int NQT = 5>
__global__ __launch_bounds__(64) void k_tazw(taz_args a) {
  __shared__ tazw_lds<CL, NQT> sh;
  if (F32 && in_lob(blockIdx.y, a.lob0, a.lob1))
    tazw_body<CL, F32, NQT, F32>(a, blockIdx.y, blockIdx.x, sh);
  else
    tazw_body<CL, F32, NQT, false>(a, blockIdx.y, blockIdx.x, sh);
}

// the FP32 bulk mode's single-disk sweep (TAZW_CL32 cells per lane)
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_tazw32(taz_args a) {
  __shared__ tazw_lds<TAZW_CL32, 5> sh;
  if (in_lob(blockIdx.y, a.lob0, a.lob1))
    tazw_body<TAZW_CL32, true, 5, true>(a, blockIdx.y, blockIdx.x, sh);
  else
    tazw_body<TAZW_CL32, true, 5, false>(a, blockIdx.y, blockIdx.x, sh);
}

// ------------------------------------------------------------ k_fin ----
// vrad recovery (Kernels.cu:1200-1205), the third damping pass, the
// temperature and, for the next substep, ConditionCFLKernel1D/2D1/2D2 of
// this ring (vrad of ring i+1 is recomputed here, it is needed by
// Kernels.cu:1302).
struct fin_args {
  const double *xd, *xvt, *xrmp, *xrmm, *press;
  double *dens, *vr, *vt, *temp;
  // adiabatic decks: the transported energy in, Energy out; the sound speed
  // field of the CFL pass (the one of the DT's start, SourceEuler.cu:208),
  // recomputed from the final state when this substep closes the DT
  const double *xe;
  double *energy, *cs;
  double gamma;
  int adiabatic, cs_refresh;
  const double *sgacc;  // self-gravity: ring means of g_r for vtheta0 (else NULL)
  // FP32 bulk mode: the fields of k_taz stored in FP32 (fp32_store_bg) with
  // the backgrounds sbar (Sigma) and vmoy (vtheta, read before this pass
  // overwrites it), and the next substep's Sigma background
  const float *fxd, *fxvt, *fxrmp, *fxrmm;
  const double *sbar;
  double *sbar_next;
  double *vmoy, *newdt;
  double *fpart;  // FP32 mode: per-ring compensated star-force partials (4 per ring), else NULL
  int *crash;     // set to 1 when a stored density is negative or not finite
  const double *nu;
  fg_dev_tables t;
  double dt, omega_frame;
  int nr, ns, evanescent, fast;
  int lob0, lob1;  // FP32 storage: rings stored as float pairs
};


// Neumaier-compensated sums (value s, correction c) of the FP32 mode's
// star force: per thread in cell order, then a fixed shuffle tree per
// wave and the waves in order
__device__ __forceinline__ void kadd(double &s, double &c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}
__device__ __forceinline__ void kmerge(double &s, double &c, double s2, double c2) {
  kadd(s, c, s2);
  c += c2;
}
__device__ __forceinline__ void kwave(double &s, double &c) {
  for (int off = 32; off > 0; off >>= 1) {
    const double s2 = __shfl_down(s, off), c2 = __shfl_down(c, off);
    kmerge(s, c, s2, c2);
  }
}

// disk force on the star from the per-ring partials, in a fixed order: each
// thread merges a contiguous run of rings, then a shuffle tree per wave,
// then the waves in order.  Called by every thread of a 256-thread block.
__device__ __forceinline__ void force_rings_final(const double *fpart, int nr, double *out2) {
  __shared__ double ws[4][4];
  const int tid = threadIdx.x, per = (nr + 255) / 256;
  double sx = 0.0, cx = 0.0, sy = 0.0, cy = 0.0;
  for (int q = 0; q < per; q++) {
    const int i = tid * per + q;
    if (i < nr) {
      kmerge(sx, cx, fpart[4 * i], fpart[4 * i + 1]);
      kmerge(sy, cy, fpart[4 * i + 2], fpart[4 * i + 3]);
    }
  }
  kwave(sx, cx);
  kwave(sy, cy);
  if ((tid & 63) == 0) {
    ws[tid >> 6][0] = sx;
    ws[tid >> 6][1] = cx;
    ws[tid >> 6][2] = sy;
    ws[tid >> 6][3] = cy;
  }
  __syncthreads();
  if (tid == 0) {
    double a = ws[0][0], b = ws[0][1], c = ws[0][2], d = ws[0][3];
    for (int w = 1; w < 4; w++) {
      kmerge(a, b, ws[w][0], ws[w][1]);
      kmerge(c, d, ws[w][2], ws[w][3]);
    }
    out2[0] = a + b;
    out2[1] = c + d;
  }
}

// FP32 mode: the star force from the per-ring partials, then the CFL count
// (ConditionCFLKernel2D2 + 2D3, Kernels.cu:1331-1359) in the same launch.
// Up to 1024 rings a thread's force partials (rings tid * per + q, merged
// in q order as force_rings_final does) are loaded together with its CFL
// inputs, so the block waits for memory once.
// The launch's sequence number goes last into slot FG_FUSED_SEQ_SLOT of the
// result block out (behind a system-scope fence: every result above is
// visible to the host first; thread 0 and the lanes that wrote them are in
// wave 0).
__device__ __forceinline__ void seq_store(double *out, double seq) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    out[FG_FUSED_SEQ_SLOT] = seq;
  }
}

__global__ __launch_bounds__(256) void k_force_rings_cfl(const double *fpart, int nr, double *out2, double *cfl,
                                                         double *dt1d, const double *vmoy, const double *newdt,
                                                         const double *invRmed, double coef, double DeltaT,
                                                         double seq) {
  if (nr > 4 * 256) {
    force_rings_final(fpart, nr, out2);
    __syncthreads();
    cfl_final(cfl, dt1d, vmoy, newdt, invRmed, coef, DeltaT, nr);
    seq_store(out2, seq);
    return;
  }
  __shared__ double ws[4][4];
  const int tid = threadIdx.x, per = (nr + 255) / 256;
  double fp[4][4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = tid * per + q;
#pragma unroll
    for (int c = 0; c < 4; c++) fp[q][c] = (q < per && i < nr) ? fpart[4 * i + c] : 0.0;
  }
  const double best = cfl_ring_min(dt1d, vmoy, newdt, invRmed, coef, nr, tid, 256);
  double sx = 0.0, cx = 0.0, sy = 0.0, cy = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (q < per && tid * per + q < nr) {
      kmerge(sx, cx, fp[q][0], fp[q][1]);
      kmerge(sy, cy, fp[q][2], fp[q][3]);
    }
  }
  kwave(sx, cx);
  kwave(sy, cy);
  if ((tid & 63) == 0) {
    ws[tid >> 6][0] = sx;
    ws[tid >> 6][1] = cx;
    ws[tid >> 6][2] = sy;
    ws[tid >> 6][3] = cy;
  }
  const double m = block_min256(best);  // (its barrier also orders ws)
  if (tid == 0) {
    double a = ws[0][0], b = ws[0][1], c = ws[0][2], d = ws[0][3];
    for (int w = 1; w < 4; w++) {
      kmerge(a, b, ws[w][0], ws[w][1]);
      kmerge(c, d, ws[w][2], ws[w][3]);
    }
    out2[0] = a + b;
    out2[1] = c + d;
  }
  cfl_count_store(cfl, m, DeltaT);
  seq_store(out2, seq);
}

// one ring; ring holds NSEC doubles (the ring's final vtheta).  FIN_B
// cells per thread have their loads issued together (4 for the single-disk
// kernel; 1 in the batched one, which would otherwise lose a wave per SIMD)
// minimum over the workgroup's 256 threads into *out (thread 0): within
// each wave by shuffles, then the four wave minima (a minimum does not
// depend on the order of its comparisons)
__device__ __forceinline__ void ring_min_store(double best, double *red, double *out) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(best, off, 64);
    if (o < best) best = o;
  }
  const int tid = threadIdx.x;
  if ((tid & 63) == 0) red[tid >> 6] = best;
  __syncthreads();
  if (tid == 0) {
    double m = red[0];
    for (int w = 1; w < 4; w++)
      if (red[w] < m) m = red[w];
    *out = m;
  }
}

#define FIN_OV_K 16  // cells per thread of the overlapped CFL pass (rings up to 3072 cells)
// PAIRS: FP32 mode, ring i - 1, i or i + 1 stored as float pairs (its own
// instance, so that the other rings run without the residuals' loads)
template <int FIN_B, bool ADI = false, bool S32 = false, bool PAIRS = false>
__device__ __forceinline__ void fin_ring(const fin_args &a, int i, double *ring) {
  __shared__ double red[4];
  __shared__ double s_vmoy;
  const int ns = a.ns, nr = a.nr, tid = threadIdx.x;
  const fg_dev_tables &t = a.t;
  const long base = (long)i * ns;
  double lam_i = 0, vr0_i = 0, vt0_i = 0, lam_n = 0, vr0_n = 0, vt0_n = 0;
  const bool damp_i = a.evanescent && t.ev_zone[i];
  const bool damp_n = a.evanescent && (i + 1 < nr) && t.ev_zone[i + 1];
  if (damp_i) ev_consts(t, a.nu, i, a.dt, a.omega_frame, lam_i, vr0_i, vt0_i, a.sgacc);
  if (damp_n) ev_consts(t, a.nu, i + 1, a.dt, a.omega_frame, lam_n, vr0_n, vt0_n, a.sgacc);
  constexpr bool adi = ADI;  // an adiabatic deck (a.adiabatic): energy and per-cell sound speed
  double fsx = 0.0, fcx = 0.0, fsy = 0.0, fcy = 0.0;
  const double Rm = t.Rmed[i], surf = t.Surf[i];
  // star at the origin (x = y = 0, no smoothing): the cell's distance is
  // Rmed, so 1/d^3 is one value per ring (FP32 mode only; the FP64 path
  // sums the reference's per-cell expression in DeviceReduce order)
  const double inv3 = 1.0 / (Rm * Rm) / Rm;
  const bool radial = i > 0 && i < nr;  // vrad recovery reads rings i-1 and i
  bool bad = false;                      // a density below 0, infinite or NaN
  // FP32 storage backgrounds of rings i-1, i, i+1 and of vtheta on ring i
  const double sb_m = S32 && i > 0 ? a.sbar[i - 1] : 0.0, sb_i = S32 ? a.sbar[i] : 0.0;
  const double vb_i = S32 ? a.vmoy[i] : 0.0;
  double vsum = 0.0, ssum = 0.0;  // FP32 mode: this thread's sums of vtheta and Sigma
  // FP32 storage: Sigma of rings i - 1, i and i + 1 stored as float pairs
  const bool lob_m = PAIRS && in_lob(i - 1, a.lob0, a.lob1), lob_i = PAIRS && in_lob(i, a.lob0, a.lob1);
  const bool lob_p = PAIRS && in_lob(i + 1, a.lob0, a.lob1);
  const long S2 = (long)(nr + 1) * ns;
  // FIN_B cells per thread per batch: all their loads first, then the math
  for (int j0 = tid; j0 < ns; j0 += FIN_B * 256) {
    double lrp[FIN_B], lrm[FIN_B], ld[FIN_B], ldm[FIN_B], lvt[FIN_B], lp[FIN_B], lcs[FIN_B], lsn[FIN_B];
#pragma unroll
    for (int u = 0; u < FIN_B; u++) {
      const int j = j0 + u * 256;
      lrp[u] = lrm[u] = ld[u] = ldm[u] = lvt[u] = lp[u] = lcs[u] = lsn[u] = 0.0;
      if (j < ns) {
        const long l = base + j;
        if constexpr (S32) {
          lcs[u] = t.cos_cell[j];
          lsn[u] = t.sin_cell[j];
          float hdm = 0.f, ydm = 0.f, yd = 0.f;
          if (radial) {
            lrp[u] = (double)a.fxrmp[l - ns];
            lrm[u] = (double)a.fxrmm[l];
            hdm = a.fxd[l - ns];
          }
          const float hd = a.fxd[l];
          lvt[u] = vb_i + (double)a.fxvt[l];
          if (radial && lob_m) ydm = a.fxd[S2 + l - ns];  // Sigma residuals of float-pair rings
          if (lob_i) yd = a.fxd[S2 + l];
          if (radial) ldm[u] = rec32(sb_m, hdm, ydm);
          ld[u] = rec32(sb_i, hd, yd);
        } else {
          if (radial) {
            lrp[u] = a.xrmp[l - ns];
            lrm[u] = a.xrmm[l];
            ldm[u] = a.xd[l - ns];
          }
          ld[u] = a.xd[l];
          lvt[u] = a.xvt[l];
        }
        if (adi)
          lp[u] = a.xe[l];
        else if (a.temp)
          lp[u] = a.press[l];
      }
    }
#pragma unroll
  for (int u = 0; u < FIN_B; u++) {
    const int j = j0 + u * 256;
    if (j >= ns) break;
    const long l = base + j;
    // ComputeVelocities, Kernels.cu:1200-1205 (0 on ring 0 and the ghost ring)
    double vr = radial ? (lrp[u] + lrm[u]) / (ld[u] + ldm[u] + 1e-20) : 0.0, vt = lvt[u], s = ld[u];
    if (damp_i) {
      vr = damp_toward(vr, lam_i, vr0_i);
      vt = damp_toward(vt, lam_i, vt0_i);
      s = damp_toward(s, lam_i, t.SigmaMed[i]);
    }
    a.dens[l] = s;
    bad |= !(s >= 0.0 && s < HUGE_VAL);
    a.vr[l] = vr;
    a.vt[l] = vt;
    ring[j] = vt;
    if constexpr (S32) {
      vsum += vt;
      ssum += s;
      const double cm = surf * s;
      kadd(fsx, fcx, cm * (Rm * lcs[u]) * inv3);
      kadd(fsy, fcy, cm * (Rm * lsn[u]) * inv3);
    }
    if (adi) {
      double e = lp[u];
      if (damp_i) e = damp_toward(e, lam_i, t.EnergyMed[i]);
      a.energy[l] = e;
      bad |= !(e >= 0.0 && e < HUGE_VAL);
      if (a.temp) a.temp[l] = (a.gamma - 1.0) * e / s;              // Kernels.cu:211
      if (a.cs_refresh) a.cs[l] = sqrt(a.gamma * (a.gamma - 1.0) * e / s);  // Kernels.cu:181
    } else if (a.temp) {
      a.temp[l] = lp[u] / s;
    }
    if (i == nr - 1) {  // ghost ring after transport: all zero
      a.dens[l + ns] = 0.0;
      a.vr[l + ns] = 0.0;
      a.vt[l + ns] = 0.0;
      if (adi) a.energy[l + ns] = 0.0;
    }
  }
  }
  if (bad) *a.crash = 1;
  __syncthreads();
  // ConditionCFLKernel2D1 + 2D2 for ring i (Kernels.cu:1284-1329)
  const double dxrad = t.Rsup[i] - t.Rinf[i];
  const double dxtheta = t.dxtheta_cfl[i];
  const double mindx = (dxtheta < dxrad) ? dxtheta : dxrad;
  const double cs = t.cs_iso[i];
  const bool radial_n = i + 1 < nr;  // vrad of ring i+1 (0 on the ghost ring)
  double best = 1e30;
  double vm;
  if constexpr (S32) {
    // FP32 mode, which does not replay the reference's summation order: the
    // ring mean of vtheta, the gravity of the ring's cells on the star
    // (Kernels.cu:703-727 at x = y = 0: the star sits at the origin with no
    // smoothing, so 1/d^3 is one value per ring) and the ring's Sigma sum
    // (the next substep's background) come from the per-thread partials of
    // the pass above, reduced per wave by a shuffle tree and over the waves
    // in order; the force as one compensated partial per ring
    __shared__ double fw[4][6];
    kwave(fsx, fcx);
    kwave(fsy, fcy);
    for (int off = 32; off > 0; off >>= 1) {
      vsum += __shfl_down(vsum, off);
      ssum += __shfl_down(ssum, off);
    }
    if ((tid & 63) == 0) {
      double *w = fw[tid >> 6];
      w[0] = fsx;
      w[1] = fcx;
      w[2] = fsy;
      w[3] = fcy;
      w[4] = vsum;
      w[5] = ssum;
    }
    __syncthreads();
    vm = (((fw[0][4] + fw[1][4]) + fw[2][4]) + fw[3][4]) / (double)ns;
    if (tid == 0) {
      double sx = fw[0][0], cx = fw[0][1], sy = fw[0][2], cy = fw[0][3];
      for (int w = 1; w < 4; w++) {
        kmerge(sx, cx, fw[w][0], fw[w][1]);
        kmerge(sy, cy, fw[w][2], fw[w][3]);
      }
      a.fpart[4 * i] = sx;
      a.fpart[4 * i + 1] = cx;
      a.fpart[4 * i + 2] = sy;
      a.fpart[4 * i + 3] = cy;
      a.sbar_next[i] = (((fw[0][5] + fw[1][5]) + fw[2][5]) + fw[3][5]) / (double)ns;
      a.vmoy[i] = vm;
    }
    if (i == 0) return;
  } else {
  if (tid == 0) {
    double m = seq_sum_lds(ring, ns);
    m /= (double)ns;
    a.vmoy[i] = m;
    s_vmoy = m;
  }
  __syncthreads();
  if (i == 0) return;
  vm = s_vmoy;
  }
  const recip64 rdr(dxrad), rdt(dxtheta);
  double smax = -1.0;
  for (int j0 = tid; j0 < ns; j0 += FIN_B * 256) {
    double pvr[FIN_B], prp[FIN_B], prm[FIN_B], pd0[FIN_B], pd1[FIN_B], pcs[FIN_B];
#pragma unroll
    for (int u = 0; u < FIN_B; u++) {
      const int j = j0 + u * 256;
      pvr[u] = prp[u] = prm[u] = pd0[u] = pd1[u] = pcs[u] = 0.0;
      if (j < ns) {
        const long l = base + j;
        pvr[u] = a.vr[l];  // written above by this thread
        if (adi) pcs[u] = a.cs[l];
        if (radial_n) {
          if constexpr (S32) {
            const double sb_p = a.sbar[i + 1];
            prp[u] = (double)a.fxrmp[l];
            prm[u] = (double)a.fxrmm[l + ns];
            const float hd1 = a.fxd[l + ns], hd0 = a.fxd[l];
            float yd1 = 0.f, yd0 = 0.f;
            if (lob_i) yd0 = a.fxd[S2 + l];  // Sigma residuals of float-pair rings
            if (lob_p) yd1 = a.fxd[S2 + l + ns];
            pd1[u] = rec32(sb_p, hd1, yd1);
            pd0[u] = rec32(sb_i, hd0, yd0);
          } else {
            prp[u] = a.xrmp[l];
            prm[u] = a.xrmm[l + ns];
            pd1[u] = a.xd[l + ns];
            pd0[u] = a.xd[l];
          }
        }
      }
    }
#pragma unroll
  for (int u = 0; u < FIN_B; u++) {
    const int j = j0 + u * 256;
    if (j >= ns) break;
    double vr = pvr[u];
    double vrn = radial_n ? (prp[u] + prm[u]) / (pd1[u] + pd0[u] + 1e-20) : 0.0;
    if (damp_n) vrn = damp_toward(vrn, lam_n, vr0_n);
    double vres = a.fast ? (ring[j] - vm) : ring[j];
    double invdt1 = (adi ? pcs[u] : cs) / mindx;
    double invdt2 = rdr.div(fabs(vr));
    double invdt3 = rdt.div(fabs(vres));
    double dvr = vrn - vr;
    double dvt = ring[j + 1 == ns ? 0 : j + 1] - ring[j];
    dvr = (dvr >= 0.0) ? 1e-10 : -dvr;
    dvt = (dvt >= 0.0) ? 1e-10 : -dvt;
    double x = rdr.div(dvr), y = rdt.div(dvt);
    double invdt4 = (y > x) ? y : x;
    invdt4 *= 4.0 * 1.41 * 1.41;
    const double sq = invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4;
    if (sq > smax) smax = sq;
  }
  }
  // the smallest rounded 0.5 / sqrt(sq) is the one at the largest sq (both
  // roundings are monotone; a NaN sq is skipped by both comparisons)
  if (smax >= 0.0) {
    const double dtc = 0.5 / sqrt(smax);
    if (dtc < best) best = dtc;
  }
  ring_min_store(best, red, a.newdt + i);
}

// FP64 rings of up to FIN_OV_K * 192 cells.  The ring mean of vtheta is a
// left-to-right sum in one lane (3072 dependent adds at 1024 x 3072), so it
// starts at once: wave 0 loads the ring's transported vtheta, damps it as
// the pass below does, stages it in LDS and lane 0 adds it up, while waves
// 1-3 do everything else in one pass over their cells: ComputeVelocities
// (Kernels.cu:1200-1205) with the energy, sound speed and temperature of an
// adiabatic deck, and every term of the cell's CFL time step
// (Kernels.cu:1284-1329) that does not need the mean.  Only the
// vtheta-residual and shear terms wait for the workgroup barrier.
// The ring minimum of dt = 0.5 / sqrt(S) over the cells is taken as one
// quotient at the largest S: rounded sqrt and rounded division are both
// monotone, so 0.5 / sqrt(max S) is the smallest rounded dt bit for bit
// (a NaN S is skipped by both comparisons; the 1e30 start caps both).
template <bool ADI>
__device__ __forceinline__ void fin_ring_ov(const fin_args &a, int i, double *ring) {
  constexpr int U = 2;  // cells per thread whose loads are issued together
  __shared__ double red[4];
  __shared__ double s_vmoy;
  const int ns = a.ns, nr = a.nr, tid = threadIdx.x;
  const fg_dev_tables &t = a.t;
  const long base = (long)i * ns;
  double lam_i = 0, vr0_i = 0, vt0_i = 0;
  const bool damp_i = a.evanescent && t.ev_zone[i];
  if (damp_i) ev_consts(t, a.nu, i, a.dt, a.omega_frame, lam_i, vr0_i, vt0_i, a.sgacc);
  const bool cfl = i > 0;  // ring 0 has no CFL term (Kernels.cu:1286)
  double A[FIN_OV_K], X[FIN_OV_K];
  if (tid < 64) {
    // the damped vtheta of every cell, which the pass below also stores
    for (int j0 = 0; j0 < ns; j0 += 64 * 16) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int j = j0 + u * 64 + tid;
        v[u] = j < ns ? a.xvt[base + j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int j = j0 + u * 64 + tid;
        if (j < ns) ring[j] = damp_i ? damp_toward(v[u], lam_i, vt0_i) : v[u];
      }
    }
    // one wave: its LDS stores complete in order before lane 0's loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (tid == 0) {
      double m = seq_sum_lds(ring, ns);
      m /= (double)ns;
      a.vmoy[i] = m;
      s_vmoy = m;
    }
  } else {
    double lam_n = 0, vr0_n = 0, vt0_n = 0;
    const bool damp_n = a.evanescent && (i + 1 < nr) && t.ev_zone[i + 1];
    if (damp_n) ev_consts(t, a.nu, i + 1, a.dt, a.omega_frame, lam_n, vr0_n, vt0_n, a.sgacc);
    const bool radial = i > 0;          // vrad recovery reads rings i-1 and i
    const bool radial_n = i + 1 < nr;  // vrad of ring i+1 (0 on the ghost ring)
    const double dxrad = t.Rsup[i] - t.Rinf[i];
    const double dxtheta = t.dxtheta_cfl[i];
    const double mindx = (dxtheta < dxrad) ? dxtheta : dxrad;
    const double invdt1r = t.cs_iso[i] / mindx;
    const recip64 rdr(dxrad);
    bool bad = false;
    const int t0 = tid - 64;
#pragma unroll
    for (int u0 = 0; u0 < FIN_OV_K; u0 += U) {
      double lrp[U], lrm[U], ld[U], ldm[U], lvt[U], lp[U], prp[U], prm[U], pd1[U], pcs[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = t0 + (u0 + u) * 192;
        lrp[u] = lrm[u] = ld[u] = ldm[u] = lvt[u] = lp[u] = prp[u] = prm[u] = pd1[u] = pcs[u] = 0.0;
        if (j < ns) {
          const long l = base + j;
          if (radial) {
            lrp[u] = a.xrmp[l - ns];
            lrm[u] = a.xrmm[l];
            ldm[u] = a.xd[l - ns];
          }
          ld[u] = a.xd[l];
          lvt[u] = a.xvt[l];
          if (ADI)
            lp[u] = a.xe[l];
          else if (a.temp)
            lp[u] = a.press[l];
          if (cfl && radial_n) {
            prp[u] = a.xrmp[l];
            prm[u] = a.xrmm[l + ns];
            pd1[u] = a.xd[l + ns];
          }
          if (ADI && cfl && !a.cs_refresh) pcs[u] = a.cs[l];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int j = t0 + (u0 + u) * 192;
        if (j >= ns) break;
        const long l = base + j;
        // ComputeVelocities, Kernels.cu:1200-1205 (0 on ring 0)
        double vr = radial ? (lrp[u] + lrm[u]) / (ld[u] + ldm[u] + 1e-20) : 0.0, vt = lvt[u], s = ld[u];
        if (damp_i) {
          vr = damp_toward(vr, lam_i, vr0_i);
          vt = damp_toward(vt, lam_i, vt0_i);
          s = damp_toward(s, lam_i, t.SigmaMed[i]);
        }
        a.dens[l] = s;
        bad |= !(s >= 0.0 && s < HUGE_VAL);
        a.vr[l] = vr;
        a.vt[l] = vt;
        double csv = pcs[u];
        if (ADI) {
          double e = lp[u];
          if (damp_i) e = damp_toward(e, lam_i, t.EnergyMed[i]);
          a.energy[l] = e;
          bad |= !(e >= 0.0 && e < HUGE_VAL);
          if (a.temp) a.temp[l] = (a.gamma - 1.0) * e / s;  // Kernels.cu:211
          if (a.cs_refresh) {
            csv = sqrt(a.gamma * (a.gamma - 1.0) * e / s);  // Kernels.cu:181
            a.cs[l] = csv;
          }
        } else if (a.temp) {
          a.temp[l] = lp[u] / s;
        }
        if (i == nr - 1) {  // ghost ring after transport: all zero
          a.dens[l + ns] = 0.0;
          a.vr[l + ns] = 0.0;
          a.vt[l + ns] = 0.0;
          if (ADI) a.energy[l + ns] = 0.0;
        }
        if (cfl) {
          double vrn = radial_n ? (prp[u] + prm[u]) / (pd1[u] + ld[u] + 1e-20) : 0.0;
          if (damp_n) vrn = damp_toward(vrn, lam_n, vr0_n);
          const double invdt1 = ADI ? csv / mindx : invdt1r;
          const double invdt2 = rdr.div(fabs(vr));
          double dvr = vrn - vr;
          dvr = (dvr >= 0.0) ? 1e-10 : -dvr;
          A[u0 + u] = invdt1 * invdt1 + invdt2 * invdt2;
          X[u0 + u] = rdr.div(dvr);
        }
      }
    }
    if (bad) *a.crash = 1;
  }
  __syncthreads();
  if (i == 0) return;
  double best = 1e30;
  if (tid >= 64) {
    const double vm = s_vmoy;
    const recip64 rdt(t.dxtheta_cfl[i]);
    double smax = -1.0;
#pragma unroll
    for (int u = 0; u < FIN_OV_K; u++) {
      const int j = tid - 64 + u * 192;
      if (j >= ns) break;
      const double vres = a.fast ? (ring[j] - vm) : ring[j];
      const double invdt3 = rdt.div(fabs(vres));
      double dvt = ring[j + 1 == ns ? 0 : j + 1] - ring[j];
      dvt = (dvt >= 0.0) ? 1e-10 : -dvt;
      const double y = rdt.div(dvt);
      double invdt4 = (y > X[u]) ? y : X[u];
      invdt4 *= 4.0 * 1.41 * 1.41;
      const double sq = (A[u] + invdt3 * invdt3) + invdt4 * invdt4;
      if (sq > smax) smax = sq;
    }
    if (smax >= 0.0) {
      const double dtc = 0.5 / sqrt(smax);
      if (dtc < best) best = dtc;
    }
  }
  ring_min_store(best, red, a.newdt + i);
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Each XCD gets a
// contiguous band of rings, so neighbouring rings, which read each other's
// momenta and densities, are processed side by side under one L2.
#define FG_NXCD 8
// ADI: adiabatic deck; S32: FP32 bulk mode (FP32 storage of k_taz's fields)
template <bool ADI, bool S32 = false>
__global__ __launch_bounds__(256) void k_fin(fin_args a) {
  extern __shared__ double lds_fin[];
  const int band = (a.nr + FG_NXCD - 1) / FG_NXCD;
  const int i = (int)(blockIdx.x % FG_NXCD) * band + (int)(blockIdx.x / FG_NXCD);
  if (i >= a.nr) return;
  if constexpr (S32) {
    if (in_lob(i, a.lob0 - 1, a.lob1 + 1))
      fin_ring<2, ADI, true, true>(a, i, lds_fin);
    else
      fin_ring<2, ADI, true, false>(a, i, lds_fin);
  } else if (a.ns <= 192 * FIN_OV_K) {
    fin_ring_ov<ADI>(a, i, lds_fin);
  } else {
    fin_ring<4, ADI, false>(a, i, lds_fin);
  }
}

// ================================================== FP32 bulk mode ====
// Same passes as above with the arithmetic in single precision.  The state
// (Sigma, vrad, vtheta) and the transported quantities between passes stay
// FP64; every FP32 result enters them as an increment added in FP64, so an
// update smaller than one FP32 ulp of the field is not lost and the fields
// do not random-walk by an ulp per substep.  Quantities whose FP32 rounding
// would cancel catastrophically are kept in FP64:
//   - star gravity against the centrifugal term in the radial momentum
//     equation (both ~1/r^2, their sum ~h^2/r^2): formed in FP64 per cell
//     from per-ring values and the cell's vtheta offset;
//   - vtheta is staged in LDS as the offset from a per-ring reference
//     (the ring mean of the previous substep), a few 1e-3 of |vtheta|;
//   - cell-planet separations are differenced in FP64;
//   - the FARGO ring mean, Ntilde/Nshift and the uniform residual, the CFL
//     count and the disk force are the FP64 code of the exact path.
// The tile height th (rings per tile: 8, 16 or 32) is chosen per grid size
// and dispatched to a template instance; the LDS tiles are carved from dynamic
// shared memory: six (th + 6) x HC float arrays and two th x TW
// increment arrays.
#define SRC32_MAX_TH 32
#define SRC32_TPB 512  // threads per FP32 source tile
#define SRC32_PER_THREAD (((SRC32_MAX_TH + HALO_R + HALO_B) * HC + SRC32_TPB - 1) / SRC32_TPB)

static size_t src32_lds_bytes(int th) { return (size_t)(6 * (th + HALO_R + HALO_B) * HC + 2 * th * TW) * sizeof(float); }

struct src32_args {
  const double *dens, *vr, *vt;
  const double *vref;  // per-ring reference vtheta
  const double *sbar;  // per-ring background of Sigma
  float *odens, *ovr, *ovt;  // FP32 storage (fp32_store_bg)
  double *opress;
  const double *nu;
  fg_dev_tables t;
  fg_planets_arg P;
  double dt, omega_frame, domega;
  int nr, ns, indirect, evanescent, corotating;
  int lob0, lob1;  // FP32 storage: rings stored as float pairs
};

// planet part of the potential at one cell (the star's -1/r is applied per
// ring by the caller); Kernels.cu:1234-1252 with the separation in FP64
__device__ __forceinline__ float cell_potential_planets_f(const fg_planets_arg &P, double Rm, double cosp,
                                                          double sinp, int indirect) {
  const double x = Rm * cosp, y = Rm * sinp;
  float acc = 0.0f;
  for (int k = 0; k < P.nb; k++) {
    const float dx = (float)(x - P.x[k]), dy = (float)(y - P.y[k]);
    const float m = (float)P.m[k];
    float pv = -m * __builtin_amdgcn_rsqf(dx * dx + dy * dy + (float)P.smooth[k]);
    if (indirect) pv += m * (float)P.invd3[k] * (float)(x * P.x[k] + y * P.y[k]);
    acc += pv;
    if (k == 0) acc -= (float)(P.ix * x + P.iy * y);
  }
  return acc;
}

// CorrectVtheta and the first damping pass of one cell's state (FP64);
// rt = the ring's tables (RT_* layout, one value per ring), ev = its
// damping constants (lambda, vrad0, vtheta0)
__device__ __forceinline__ void src32_adjust(const src32_args &a, int i, double rmed, double evz, double smed,
                                             const double *ev, double &s, double &vr, double &vt) {
  if (i < a.nr) {
    if (a.corotating) vt = vt - a.domega * rmed;
    if (a.evanescent && evz != 0.0) {
      vr = damp_toward(vr, ev[0], ev[1]);
      vt = damp_toward(vt, ev[0], ev[2]);
      s = damp_toward(s, ev[0], smed);
    }
  }
}

#define SRC32_OWN_MAX (SRC32_MAX_TH * TW / SRC32_TPB)  // own cells per thread

// staged position (row r, column cc of the TH+6 x HC stage) of halo cell h
// of a tile of TH rows: the HALO_R rows above, the HALO_B rows below, then
// the HALO_C columns on either side of each tile row
template <int TH_>
__device__ __forceinline__ void halo_cell(int h, int &r, int &cc) {
  constexpr int NTOP = HALO_R * HC, NBOT = HALO_B * HC;
  if (h < NTOP) {
    r = h / HC;
    cc = h % HC;
  } else if (h < NTOP + NBOT) {
    h -= NTOP;
    r = HALO_R + TH_ + h / HC;
    cc = h % HC;
  } else {
    h -= NTOP + NBOT;
    const int c = h % (2 * HALO_C);
    r = HALO_R + h / (2 * HALO_C);
    cc = c < HALO_C ? c : c + TW;
  }
}

// the damping constants of the tile's rings are computed once per ring
// into LDS
// ROWS (rings per tile) is a template parameter so that the LDS array offsets
// and the per-thread trip counts are constants
template <int ROWS>
__device__ __forceinline__ void src32_tile(const src32_args &a, int bx, int by, float *lds) {
  constexpr int th = ROWS, hr = th + HALO_R + HALO_B;
  float(*sS)[HC] = (float(*)[HC])(lds);           // Sigma after the first damping
  float(*sA)[HC] = (float(*)[HC])(lds + hr * HC);  // pressure, then VradInt
  float(*sB)[HC] = (float(*)[HC])(lds + 2 * hr * HC);  // planet potential, then VthetaInt - vref
  float(*sC)[HC] = (float(*)[HC])(lds + 3 * hr * HC);  // vtheta - vref, then TAURR, then Qr
  float(*sD)[HC] = (float(*)[HC])(lds + 4 * hr * HC);  // vrad, then TAUPP, then Qtheta
  float(*sE)[HC] = (float(*)[HC])(lds + 5 * hr * HC);  // TAURP
  float(*incR)[TW] = (float(*)[TW])(lds + 6 * hr * HC);  // own-cell velocity increments
  float(*incT)[TW] = (float(*)[TW])(lds + 6 * hr * HC + th * TW);
  __shared__ double evc[SRC32_MAX_TH + HALO_R + HALO_B][3];  // damping constants per tile row
  const int nr = a.nr, ns = a.ns;
  const int i0 = by * th, j0 = bx * TW;
  const int tid = threadIdx.x;
  const fg_dev_tables &t = a.t;
  const double dt = a.dt;
  const float dtf = (float)dt;
#define GI(r) (i0 - HALO_R + (r))
#define GJ(c) wrap_col(j0 - HALO_C + (c), ns)
  // per-ring tables of the tile's rows as in k_src, plus the reference
  // vtheta (0 outside the disk), and the potential's cos/sin per column
  constexpr int RT_VREF = RT_N, RT_SBAR = RT_N + 1, MAXR = SRC32_MAX_TH + HALO_R + HALO_B;
  __shared__ double rt[RT_N + 2][MAXR];
  __shared__ double ct[2][HC];

  // ---- S0: load, CorrectVtheta, damping #1 (FP64), pressure, potential ----
  // all of a thread's loads are issued before any of them is used.  The
  // tile's own cells are loaded in the mapping of the final stage (S5) and
  // their adjusted FP64 state stays in registers until then; the halo
  // cells (rows above and below, columns on either side) separately
  constexpr int NH = hr * HC - th * TW, HPT = (NH + SRC32_TPB - 1) / SRC32_TPB;
  double os[SRC32_OWN_MAX], ovr[SRC32_OWN_MAX], ovt[SRC32_OWN_MAX];
  {
    double ls[HPT], lvr[HPT], lvt[HPT];
#pragma unroll
    for (int k = 0; k < SRC32_OWN_MAX; k++) {
      const int q = tid + k * SRC32_TPB;
      os[k] = ovr[k] = ovt[k] = 0.0;
      if (q < th * TW) {
        const int r = q / TW + HALO_R, cc = q % TW + HALO_C, i = GI(r);
        if (i >= 0 && i <= nr) {
          const long l = (long)i * ns + GJ(cc);
          os[k] = a.dens[l];
          ovr[k] = a.vr[l];
          ovt[k] = a.vt[l];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < HPT; k++) {
      const int h = tid + k * SRC32_TPB;
      ls[k] = lvr[k] = lvt[k] = 0.0;
      if (h < NH) {
        int r, cc;
        halo_cell<th>(h, r, cc);
        const int i = GI(r);
        if (i >= 0 && i <= nr) {
          const long l = (long)i * ns + GJ(cc);
          ls[k] = a.dens[l];
          lvr[k] = a.vr[l];
          lvt[k] = a.vt[l];
        }
      }
    }
    // tables staged while the loads above are in flight: wave w loads
    // tables w, w + 4, w + 8 (wave 0 also the ev_zone flags, wave 1 the
    // reference vtheta) of rows ln and ln + 64
    {
      const int w = tid >> 6, ln = tid & 63;
      for (int r = ln; w < 4 && r < hr; r += 64) {
        const int gi = GI(r), i = min(max(gi, 0), nr - 1);
        const double v0 = ring_table(t, a.nu, w)[i], v1 = ring_table(t, a.nu, w + 4)[i];
        const double v2 = ring_table(t, a.nu, w + 8)[i];
        const int ez = t.ev_zone[i];
        const double vref = (gi >= 0 && gi < nr) ? a.vref[gi] : 0.0;
        const double sbar = (gi >= 0 && gi < nr) ? a.sbar[gi] : 0.0;
        rt[w][r] = v0;
        rt[w + 4][r] = v1;
        rt[w + 8][r] = v2;
        if (w == 0) rt[RT_EVZ][r] = (double)ez;
        if (w == 1) rt[RT_VREF][r] = vref;
        if (w == 2) rt[RT_SBAR][r] = sbar;
      }
      for (int c = tid; c < HC; c += SRC32_TPB) {
        const int j = GJ(c);
        ct[0][c] = t.cos_pot[j];
        ct[1][c] = t.sin_pot[j];
      }
    }
    // damping constants of the tile's rings, once per ring (Kernels.cu:978-1004)
    if (a.evanescent) {
      for (int r = tid; r < hr; r += SRC32_TPB) {
        const int i = GI(r);
        if (i >= 0 && i < nr && t.ev_zone[i]) ev_consts(t, a.nu, i, dt, a.omega_frame, evc[r][0], evc[r][1], evc[r][2]);
      }
    }
    __syncthreads();
    // CorrectVtheta and damping #1 of a staged cell (the FP64 values in
    // place), then its FP32 stage values
    auto stage = [&](int r, int cc, double &sd, double &vrd, double &vtd) {
      const int i = GI(r);
      float s = 0.f, vr = 0.f, vt = 0.f, p = 0.f, phi = 0.f;
      if (i >= 0 && i <= nr) {
        src32_adjust(a, i, rt[RT_MED][r], rt[RT_EVZ][r], rt[RT_SMED][r], evc[r], sd, vrd, vtd);
        s = (float)sd;
        vr = (float)vrd;
        vt = (float)(vtd - rt[RT_VREF][r]);
        if (i < nr) {
          const double cs = rt[RT_CS][r];
          p = (float)(sd * cs * cs);
          phi = cell_potential_planets_f(a.P, rt[RT_MED][r], ct[0][cc], ct[1][cc], a.indirect);
        }
      }
      sS[r][cc] = s;
      sA[r][cc] = p;
      sB[r][cc] = phi;
      sC[r][cc] = vt;
      sD[r][cc] = vr;
    };
#pragma unroll
    for (int k = 0; k < SRC32_OWN_MAX; k++) {
      const int q = tid + k * SRC32_TPB;
      if (q >= th * TW) break;
      stage(q / TW + HALO_R, q % TW + HALO_C, os[k], ovr[k], ovt[k]);
    }
#pragma unroll
    for (int k = 0; k < HPT; k++) {
      const int h = tid + k * SRC32_TPB;
      if (h >= NH) break;
      int r, cc;
      halo_cell<th>(h, r, cc);
      stage(r, cc, ls[k], lvr[k], lvt[k]);
    }
  }
  __syncthreads();

  // ---- S1: Substep1 (Kernels.cu:13-53) ----
  const bool star = a.P.nb > 0;  // the star's -1/r enters with the first planet
  float regA[SRC32_PER_THREAD], regB[SRC32_PER_THREAD];
  {
#pragma unroll
    for (int k = 0; k < SRC32_PER_THREAD; k++) {
      const int q = tid + k * SRC32_TPB;
      if (q >= hr * HC) break;
      int r = q / HC, cc = q - r * HC;
      int i = GI(r);
      float vri = 0.f, vti = 0.f, dvr = 0.f, dvt = 0.f;
      if (r >= 1 && cc >= 1 && cc < HC - 1 && i >= 0 && i < nr) {
        if (i > 0) {
          const float idr = (float)rt[RT_IDMED][r];
          float gradp = (sA[r][cc] - sA[r - 1][cc]) * 2.0f * rcp_f(sS[r][cc] + sS[r - 1][cc]) * idr;
          float gradphi = (sB[r][cc] - sB[r - 1][cc]) * idr;
          // centrifugal minus star gravity in FP64: vbar = ring part + offset
          double vbar = 0.5 * (rt[RT_VREF][r] + rt[RT_VREF][r - 1]) + a.omega_frame * rt[RT_INF][r];
          vbar += (double)((sC[r][cc] + sC[r][cc + 1] + sC[r - 1][cc] + sC[r - 1][cc + 1]) * 0.25f);
          double bal = vbar * vbar * rt[RT_IINF][r];
          if (star) bal -= (rt[RT_IMED][r - 1] - rt[RT_IMED][r]) * rt[RT_IDMED][r];
          dvr = dtf * (-gradp - gradphi + (float)bal);
          vri = sD[r][cc] + dvr;
        }
        const float invdxtheta = (float)rt[RT_IDXT][r];
        float gradp = (sA[r][cc] - sA[r][cc - 1]) * 2.0f * rcp_f(sS[r][cc] + sS[r][cc - 1]) * invdxtheta;
        float gradphi = (sB[r][cc] - sB[r][cc - 1]) * invdxtheta;
        dvt = dtf * ((float)rt[RT_SUPP][r] - (gradp + gradphi));
        vti = sC[r][cc] + dvt;
      }
      regA[k] = vri;
      regB[k] = vti;
      const int ro = r - HALO_R, co = cc - HALO_C;
      if (ro >= 0 && ro < th && co >= 0 && co < TW) {
        incR[ro][co] = dvr;
        incT[ro][co] = dvt;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SRC32_PER_THREAD; k++) {
      const int q = tid + k * SRC32_TPB;
      if (q >= hr * HC) break;
      int r = q / HC, cc = q - r * HC;
      sA[r][cc] = regA[k];
      sB[r][cc] = regB[k];
    }
  }
  __syncthreads();

  // ---- S2: viscous stress (Kernels.cu:733-771) ----
  const float invdphi = (float)t.invdphi;
  const float onethird = 1.0f / 3.0f;
#pragma unroll
  for (int kq = 0; kq < SRC32_PER_THREAD; kq++) {
    const int q = tid + kq * SRC32_TPB;
    if (q >= hr * HC) break;
    int r = q / HC, cc = q - r * HC;
    int i = GI(r);
    if (r < 1 || r >= hr - 1 || cc < 2 || cc >= HC - 2 || i < 0 || i >= nr) continue;
    const float iRm = (float)rt[RT_IMED][r], idRs = (float)rt[RT_IDSUP][r];
    const float vr0 = sA[r][cc], vr1 = sA[r + 1][cc];
    const float dvt = sB[r][cc + 1] - sB[r][cc];
    float drr = (vr1 - vr0) * idRs;
    float dpp = dvt * invdphi * iRm + 0.5f * (vr1 + vr0) * iRm;
    float dv = (vr1 * (float)rt[RT_SUP][r] - vr0 * (float)rt[RT_INF][r]) * idRs * iRm;
    dv += dvt * invdphi * iRm;
    const float nu2s = 2.0f * sS[r][cc] * (float)rt[RT_NU][r];
    sC[r][cc] = nu2s * (drr - onethird * dv);
    sD[r][cc] = nu2s * (dpp - onethird * dv);
    if (i > 0) {
      // d(vtheta/r)/dr with the reference part of vtheta differenced in FP64
      const double dref = (rt[RT_VREF][r] * rt[RT_IMED][r] - rt[RT_VREF][r - 1] * rt[RT_IMED][r - 1]);
      const float dom = (float)dref + (sB[r][cc] * iRm - sB[r - 1][cc] * (float)rt[RT_IMED][r - 1]);
      float drp = 0.5f * ((float)rt[RT_INF][r] * dom * (float)rt[RT_IDMED][r] +
                          (sA[r][cc] - sA[r][cc - 1]) * invdphi * (float)rt[RT_IINF][r]);
      sE[r][cc] = 0.5f * (sS[r][cc] + sS[r - 1][cc] + sS[r][cc - 1] + sS[r - 1][cc - 1]) * (float)rt[RT_NU][r] * drp;
    }
  }
  __syncthreads();

  // ---- S3: viscous velocity update (Kernels.cu:118-150) ----
#pragma unroll
  for (int kq = 0; kq < SRC32_PER_THREAD; kq++) {
    const int q = tid + kq * SRC32_TPB;
    if (q >= hr * HC) break;
    int r = q / HC, cc = q - r * HC;
    int i = GI(r);
    if (r < 2 || r >= hr - 1 || cc < 3 || cc >= HC - 3 || i < 1 || i >= nr) continue;
    const float iRm = (float)rt[RT_IMED][r];
    const int ro = r - HALO_R, co = cc - HALO_C;
    const bool own = ro >= 0 && ro < th && co >= 0 && co < TW;
    if (i < nr - 1 && r < hr - 2) {
      float d = dtf * iRm *
                (((float)rt[RT_SUP][r] * sE[r + 1][cc] - (float)rt[RT_INF][r] * sE[r][cc]) * (float)rt[RT_IDSUP][r] +
                 (sD[r][cc] - sD[r][cc - 1]) * invdphi + 0.5f * (sE[r][cc] + sE[r + 1][cc])) *
                rcp_f(0.5f * (sS[r][cc] + sS[r][cc - 1]));
      sB[r][cc] += d;
      if (own) incT[ro][co] += d;
    }
    float d = dtf * (float)rt[RT_IINF][r] *
              (((float)rt[RT_MED][r] * sC[r][cc] - (float)rt[RT_MED][r - 1] * sC[r - 1][cc]) * (float)rt[RT_IDMED][r] +
               (sE[r][cc + 1] - sE[r][cc]) * invdphi - 0.5f * (sD[r][cc] + sD[r - 1][cc])) *
              rcp_f(0.5f * (sS[r][cc] + sS[r - 1][cc]));
    sA[r][cc] += d;
    if (own) incR[ro][co] += d;
  }
  __syncthreads();

  // ---- S4: artificial pressure (Kernels.cu:409-434) ----
#pragma unroll
  for (int kq = 0; kq < SRC32_PER_THREAD; kq++) {
    const int q = tid + kq * SRC32_TPB;
    if (q >= hr * HC) break;
    int r = q / HC, cc = q - r * HC;
    int i = GI(r);
    if (r < 3 || r >= hr - 1 || cc < 3 || cc >= HC - 4 || i < 0 || i >= nr) continue;
    sC[r][cc] = art_visc_q_f(sS[r][cc], sA[r + 1][cc] - sA[r][cc]);
    sD[r][cc] = art_visc_q_f(sS[r][cc], sB[r][cc + 1] - sB[r][cc]);
  }
  __syncthreads();

  // ---- S5: velocity update ("kernel", Kernels.cu:438-479), damping #2 ----
  // (the own cells' state as adjusted by S0 is still in os, ovr, ovt)
#pragma unroll
  for (int k = 0; k < SRC32_OWN_MAX; k++) {
    const int q = tid + k * SRC32_TPB;
    if (q >= th * TW) break;
    int ro = q / TW, co = q % TW;
    int r = ro + HALO_R, cc = co + HALO_C;
    int i = GI(r), j = j0 + co;
    if (i > nr || j >= ns) continue;
    long l = (long)i * ns + j;
    double s = os[k], vr = ovr[k], vt = ovt[k];
    const double *ev = evc[r];
    if (a.opress && i < nr) a.opress[l] = s * rt[RT_CS][r] * rt[RT_CS][r];
    if (i == nr) {  // ghost ring (never read)
      a.odens[l] = 0.0f;
      a.ovr[l] = 0.0f;
      a.ovt[l] = 0.0f;
      continue;
    }
    double vrnew = 0.0;
    if (i > 0) {
      float dq = (sC[r][cc] - sC[r - 1][cc]) * 2.0f * rcp_f(sS[r][cc] + sS[r - 1][cc]) * (float)rt[RT_IDMED][r];
      vrnew = vr + (double)(incR[ro][co] - dtf * dq);
    }
    float dq = (sD[r][cc] - sD[r][cc - 1]) * 2.0f * rcp_f(sS[r][cc] + sS[r][cc - 1]) * (float)rt[RT_IDXT][r];
    double vtnew = vt + (double)(incT[ro][co] - dtf * dq);
    if (a.evanescent && rt[RT_EVZ][r] != 0.0) {
      vrnew = damp_toward(vrnew, ev[0], ev[1]);
      vtnew = damp_toward(vtnew, ev[0], ev[2]);
      s = damp_toward(s, ev[0], rt[RT_SMED][r]);
    }
    const double xs = s - rt[RT_SBAR][r], xv = vtnew - rt[RT_VREF][r];
    const float hs = (float)xs, hv = (float)vrnew, ht = (float)xv;
    a.odens[l] = hs;
    a.ovr[l] = hv;
    a.ovt[l] = ht;
    if (in_lob(i, a.lob0, a.lob1))  // Sigma as a float pair near the planets
      a.odens[(long)(nr + 1) * ns + l] = (float)(xs - (double)hs);
  }
#undef GI
#undef GJ
}

// one kernel per tile height (8, 16 or 32, the heights fg_set_tiles
// accepts): each instance needs only its own registers
template <int ROWS>
__global__ __launch_bounds__(SRC32_TPB) void k_src32(src32_args a) {
  extern __shared__ float lds_src32[];
  src32_tile<ROWS>(a, blockIdx.x, blockIdx.y, lds_src32);
}

// Radial sweep in FP32 (k_trad above): Q of each ring is formed in FP64,
// its Work value, slopes, face values and fluxes in FP32; the flux
// divergence is added to the FP64 Q.
// block index blk < NRAD: ring mean of that ring; otherwise one column
// strip of one band of a.band rings
__device__ __forceinline__ void trad32_block(const trad_args &a, int blk) {
  const int nr = a.nr, ns = a.ns;
  if (blk < nr) {
    ring_mean_wave32(a.fvt + (long)blk * ns, a.vref[blk], ns, a.vmed + blk);
    return;
  }
  const int bid = blk - nr;
  // lanes past the last sector: as in k_trad
  const int j = min((bid % a.nstrips) * TRAD_TPB + (int)threadIdx.x, ns - 1);
  const int jp = (j + 1) % ns;
  const int b0 = (bid / a.nstrips) * a.band;
  const int b1 = min(b0 + a.band, nr);
  const fg_dev_tables &t = a.t;
  const double dt = a.dt, OF = a.omega_frame;
  const float dtf = (float)dt;
  const float cflux = (float)(dt * t.dtheta);
  double Q[NQ][3] = {};
  float W[NQ][3] = {}, dq[NQ][2] = {}, F[NQ][2] = {};
  float sg[3] = {}, dqs[2] = {}, vr[3] = {};
  const int m0 = max(b0 - 2, 0);
  ring_lanes rt;
  rt.load(t, m0 - 2, nr);
  bg_lanes bg;
  bg.load(a.sbar, a.vref, rt.r0, nr);
  // Every iteration issues the same memory operations, so that the
  // compiler's wait counts are exact: ring m + 2's inputs are loaded two
  // iterations ahead (the Sigma residual too, from the ghost ring's zero
  // slots outside the float-pair band), and ring m - 2's outputs are
  // stored (the iterations before the strip's first stored ring, and the
  // residuals outside the band, write the never-read ghost-ring slots of
  // the output arrays).  A sweep has one wave per column strip and fewer
  // strips than SIMDs, so only its own prefetch hides memory latency; the
  // three register sets P0..P2 take turns (ring m in one, ring m + 2
  // loaded into the one ring m - 1 left) in a loop unrolled by three, so
  // no register has to be renamed while its load is in flight (each set
  // loads its own ring's vrad as v0, a cache hit: the previous ring loaded
  // it as its v1).
  const long S = (long)(nr + 1) * ns, ghost = (long)nr * ns;
  auto lo_at = [&](int ring, long l) { return in_lob(ring, a.lob0, a.lob1) ? S + l : S + ghost + (l - (long)ring * ns); };
  struct pset {
    float s, v0, v1, vtp, vt, sl;
  };
  auto load = [&](pset &p, int mr) {
    const int mp = min(mr, nr - 1);
    const long l = (long)mp * ns;
    p.s = a.fdens[l + j];
    p.v0 = a.fvr[l + j];
    p.v1 = a.fvr[l + ns + j];
    p.vtp = a.fvt[l + jp];
    p.vt = a.fvt[l + j];
    p.sl = a.fdens[lo_at(mp, l + j)];
  };
  pset P0, P1, P2;
  load(P0, m0);
  load(P1, m0 + 1);
  asm volatile("" ::"v"(P0.s), "v"(P0.v0), "v"(P0.v1), "v"(P0.vtp), "v"(P0.vt), "v"(P0.sl));
  asm volatile("" ::"v"(P1.s), "v"(P1.v0), "v"(P1.v1), "v"(P1.vtp), "v"(P1.vt), "v"(P1.sl));
  auto ring = [&](int m, pset &cur, pset &next2) {
    const int r = rt.at(t, m, nr);
    if (bg.r0 != rt.r0) bg.load(a.sbar, a.vref, rt.r0, nr);
    const float c_s = cur.s, c_v0 = cur.v0, c_v1 = cur.v1, c_vtp = cur.vtp, c_vt = cur.vt;
    const float c_sl = cur.sl;  // 0 outside the band
    load(next2, m + 2);
    for (int q = 0; q < NQ; q++) {
      Q[q][0] = Q[q][1];
      Q[q][1] = Q[q][2];
      W[q][0] = W[q][1];
      W[q][1] = W[q][2];
    }
    sg[0] = sg[1];
    sg[1] = sg[2];
    vr[0] = vr[1];
    vr[1] = vr[2];
    const double Rm = lane_get(rt.Rmed, r);
    if (m < nr) {
      const double vb = lane_get(bg.vb, r);
      const double s = rec32(lane_get(bg.sb, r), c_s, c_sl);
      const double v0 = (double)c_v0, v1 = (double)c_v1;
      Q[0][2] = s * v1;
      Q[1][2] = s * v0;
      Q[2][2] = s * ((vb + (double)c_vtp) + Rm * OF) * Rm;
      Q[3][2] = s * ((vb + (double)c_vt) + Rm * OF) * Rm;
      Q[4][2] = s;
      const float fs = (float)s;
      const float rs = rcp_f(fs + 1e-20f);
      for (int q = 0; q < NQ; q++) W[q][2] = (float)Q[q][2] * rs;
      sg[2] = fs;
      vr[2] = c_v0;
    } else {
      for (int q = 0; q < NQ; q++) {
        Q[q][2] = 0.0;
        W[q][2] = 0.0f;
      }
      sg[2] = 0.0f;
      vr[2] = 0.0f;
    }
    // (k = -1, the first step of the first strip: its slopes and fluxes are
    // of zero quantities and feed no stored ring)
    const int k = m - 1;
    const bool flat = (k == 0 || k >= nr - 1);
    const float idm0 = flat ? 0.f : (float)lane_get(rt.idm, r - 1), idm1 = flat ? 0.f : (float)lane_get(rt.idm, r);
    for (int q = 0; q < NQ; q++) {
      dq[q][0] = dq[q][1];
      dq[q][1] = flat ? 0.0f : vl_rad_slope_f(W[q][0], W[q][1], W[q][2], idm0, idm1);
    }
    dqs[0] = dqs[1];
    dqs[1] = flat ? 0.0f : vl_rad_slope_f(sg[0], sg[1], sg[2], idm0, idm1);
    float DS = 0.0f, QS[NQ];
    if (k > 0 && k < nr) {
      const float vk = vr[1];
      const double Rk = lane_get(rt.Rmed, r - 1);
      const float dm = (float)(Rk - lane_get(rt.Rmed, r - 2)), dp = (float)(Rm - Rk);
      DS = vl_rad_star_f(vk, dtf, sg[0], sg[1], dqs[0], dqs[1], dm, dp);
      for (int q = 0; q < NQ; q++) QS[q] = vl_rad_star_f(vk, dtf, W[q][0], W[q][1], dq[q][0], dq[q][1], dm, dp);
    } else {
      for (int q = 0; q < NQ; q++) QS[q] = 0.0f;
    }
    const float vface = (k < nr) ? vr[1] : 0.0f;
    const float cf = cflux * (float)lane_get(rt.Rinf, r - 1) * DS * vface;
    for (int q = 0; q < NQ; q++) {
      F[q][0] = F[q][1];
      F[q][1] = cf * QS[q];
    }
    {
      // ring o = k - 1 is complete: stored as FP32 offsets from its
      // backgrounds if the strip stores it (o in [b0, b1)), else into the
      // ghost ring's slots, which nothing reads
      const int o = k - 1;
      const long lo_ = (o >= b0 ? (long)o * ns : ghost) + j;
      const float isurf = (float)lane_get(rt.isurf, r - 2);
      const double sb = lane_get(bg.sb, r - 2);
      double val[NQ];
      for (int q = 0; q < NQ; q++) val[q] = Q[q][0] + (double)((F[q][0] - F[q][1]) * isurf);
      // Sigma as the reader will reconstruct it, the angular momenta's background
      const double xs = val[4] - sb;
      const float fs = (float)xs, fsl = (float)(xs - (double)fs);
      const bool pair = in_lob(o, a.lob0, a.lob1);
      const float fsr = pair ? fsl : 0.f;
      const double lb = fp32_bg_angmom(rec32(sb, fs, fsr), lane_get(bg.vb, r - 2), lane_get(rt.Rmed, r - 2), OF);
      a.fout[0][lo_] = (float)val[0];
      a.fout[1][lo_] = (float)val[1];
      a.fout[2][lo_] = (float)(val[2] - lb);
      a.fout[3][lo_] = (float)(val[3] - lb);
      a.fout[4][lo_] = fs;
      // the residual of a float-pair ring (k_taz reads it inside the band only)
      a.fout[4][S + (pair && o >= b0 ? lo_ : ghost + j)] = fsl;
    }
  };
  int m = m0;
  const int end = b1 + 2;  // steps m0 .. b1 + 1 store rings b0 .. b1 - 1
  for (; m + 2 < end; m += 3) {
    ring(m, P0, P2);
    ring(m + 1, P1, P0);
    ring(m + 2, P2, P1);
  }
  if (m < end) ring(m++, P0, P2);
  if (m < end) ring(m++, P1, P0);
}

__global__ __launch_bounds__(TRAD_TPB) void k_trad32(trad_args a) { trad32_block(a, blockIdx.x); }

// ----------------------------------------------------- entry points ----
#define PROF_KERNELS 6

static void prof_mark(fg_ctx *c) {
  if (c->prof_on && c->prof_nev < 16) (void)hipEventRecord(c->prof_ev[c->prof_nev++], c->stream);
}

// fold a completed substep's events into the totals (stream is idle)
static void prof_fold(fg_ctx *c) {
  if (!c->prof_on || !c->prof_pending) return;
  for (int k = 0; k + 1 < c->prof_nev && k < PROF_KERNELS; k++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->prof_ev[k], c->prof_ev[k + 1]) == hipSuccess) {
      c->prof_ms[k] += ms;
      c->prof_calls[k]++;
    }
  }
  c->prof_pending = 0;
  c->prof_nev = 0;
}

extern "C" fg_status fg_profile(fg_ctx *c, int enable) {
  if (!c) return FG_E_INVALID;
  FG_HIP(hipStreamSynchronize(c->stream));
  if (enable && !c->prof_ev[0])
    for (int k = 0; k < 16; k++) FG_HIP(hipEventCreate(&c->prof_ev[k]));
  for (int k = 0; k < 16; k++) {
    c->prof_ms[k] = 0.0;
    c->prof_calls[k] = 0;
  }
  c->prof_nev = 0;
  c->prof_pending = 0;
  c->prof_on = enable;
  return FG_OK;
}

extern "C" int fg_profile_read(fg_ctx *c, double *ms, long *calls, int max) {
  if (!c) return 0;
  (void)hipStreamSynchronize(c->stream);
  prof_fold(c);
  int n = max < PROF_KERNELS ? max : PROF_KERNELS;
  for (int k = 0; k < n; k++) {
    if (ms) ms[k] = c->prof_ms[k];
    if (calls) calls[k] = c->prof_calls[k];
  }
  return n;
}
// The fused path's buffers are published together, only once every
// allocation succeeded (a failed attempt frees what it got and a later call
// starts again).
// arrays in the fused pool: B (3), R (5), X (4), pressure, spare; adiabatic
// or self-gravity decks add EnergyInt, two energies and the two kicks
static int fused_pool_arrays(const fg_ctx *c) { return (c->p.adiabatic || c->p.self_gravity) ? 19 : 14; }

static fg_status fused_alloc(fg_ctx *c) {
  if (c->fu_ring) return FG_OK;
  const size_t S = (size_t)c->S;
  // per-ring star-force partials (4 per ring) and the two Sigma backgrounds
  // of the FP32 storage
  const size_t fpart_bytes = 6 * (size_t)(c->nr + 1) * sizeof(double);
  const size_t pool_bytes = (size_t)fused_pool_arrays(c) * S * sizeof(double);
  double *pool = nullptr, *fpart = nullptr, *h = nullptr, *d = nullptr;
  hipError_t e = hipMalloc((void **)&pool, pool_bytes);
  if (e == hipSuccess) e = hipMemset(pool, 0, pool_bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&fpart, fpart_bytes);
  if (e == hipSuccess) e = hipMemset(fpart, 0, fpart_bytes);
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&h, FG_FUSED_SLOTS * sizeof(double), hipHostMallocCoherent | hipHostMallocMapped);
  if (e == hipSuccess) {
    memset(h, 0, FG_FUSED_SLOTS * sizeof(double));
    e = hipHostGetDevicePointer((void **)&d, h, 0);
  }
  if (e != hipSuccess) {
    if (pool) (void)hipFree(pool);
    if (fpart) (void)hipFree(fpart);
    if (h) (void)hipHostFree(h);
    fg_set_error("fused path buffers: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? FG_E_NOMEM : FG_E_HIP;
  }
  c->fu_fpart = fpart;
  c->fu_sbar[0] = fpart + 4 * (size_t)(c->nr + 1);
  c->fu_sbar[1] = c->fu_sbar[0] + (c->nr + 1);
  c->fu_sbar_cur = 0;
  c->h_fused = h;
  c->d_fused = d;
  c->fu_ring = pool;
  return FG_OK;
}

static fused_bufs bufs_of(fg_ctx *c) {
  const size_t S = (size_t)c->S;
  double *p = c->fu_ring;
  fused_bufs b;
  b.Bd = p + 0 * S;
  b.Bvr = p + 1 * S;
  b.Bvt = p + 2 * S;
  b.Rd = p + 3 * S;
  b.Rrmp = p + 4 * S;
  b.Rrmm = p + 5 * S;
  b.Rtmp = p + 6 * S;
  b.Rtmm = p + 7 * S;
  b.Xd = p + 8 * S;
  b.Xvt = p + 9 * S;
  b.Xrmp = p + 10 * S;
  b.Xrmm = p + 11 * S;
  b.P = p + 12 * S;
  const bool more = fused_pool_arrays(c) > 14;
  b.Beint = more ? p + 14 * S : nullptr;
  b.Re = more ? p + 15 * S : nullptr;
  b.Xe = more ? p + 16 * S : nullptr;
  b.kick_r = more ? p + 17 * S : nullptr;
  b.kick_t = more ? p + 18 * S : nullptr;
  return b;
}

static fg_status fused_supported(fg_ctx *c) {
  const fg_params &p = c->p;
  if (p.open_inner || p.non_reflecting || !p.evanescent) {
    fg_set_error("fused path: deck needs the per-routine path (open or non-reflecting edges)");
    return FG_E_UNSUPPORTED;
  }
  if ((p.adiabatic || p.self_gravity) && p.precision == FG_PREC_FP32) {
    fg_set_error("fused path: the FP32 bulk mode takes locally isothermal decks without self-gravity");
    return FG_E_UNSUPPORTED;
  }
  if (p.adiabatic && c->nr < 3) {
    fg_set_error("fused path: an adiabatic deck needs NRAD >= 3");
    return FG_E_UNSUPPORTED;
  }
  if (c->ns > 4096) {
    fg_set_error("fused path: NSEC %d > 4096", c->ns);
    return FG_E_UNSUPPORTED;
  }
  return FG_OK;
}

// The host loop of DeviceReduce for the star force (dreduce_final: lanes 0
// and 1 add the B block sums of x and y in order, then +0.0 for the inner
// sum, SideEuler.cu:144) and, in waves 1-3 meanwhile, the CFL count of the
// same substep (DT1D of every ring, Kernels.cu:1331-1337, and the minimum
// with the ring minima, Kernels.cu:1337-1359; a minimum does not depend on
// the order its candidates are compared in), then the result block's
// sequence slot.  p: LDS for the 2B block sums; every thread of a
// 256-thread block calls it.
__device__ __forceinline__ void force_cfl_final(const double *partial, int B, double *out, double *cfl, double *dt1d,
                                                const double *vmoy, const double *newdt, const double *invRmed,
                                                double coef, double DeltaT, int nr, double seq, double *p) {
  __shared__ double wmin[4];
  const int tid = threadIdx.x;
  for (int k = tid; k < 2 * B; k += 256) p[k] = partial[k];
  __syncthreads();
  if (tid < 2) {
    double s = seq_sum_lds(p + tid * B, B);
    out[tid] = 0.0 + s;
  } else if (tid >= 64) {
    double best = cfl_ring_min(dt1d, vmoy, newdt, invRmed, coef, nr, tid - 64, 192);
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(best, off, 64);
      if (o < best) best = o;
    }
    if ((tid & 63) == 0) wmin[tid >> 6] = best;
  }
  __syncthreads();
  if (tid == 0) {
    double newdt_ = wmin[1];
    if (wmin[2] < newdt_) newdt_ = wmin[2];
    if (wmin[3] < newdt_) newdt_ = wmin[3];
    if (DeltaT < newdt_) newdt_ = DeltaT;
    cfl[0] = (double)(int)(ceil(DeltaT / newdt_));
  }
  seq_store(out, seq);
}

// the DeviceReduce-order partials of the star force (dreduce_block)
template <int CH>
__global__ __launch_bounds__(RED_TPB) void k_dreduce_partial_f(force_src src, long n, int B, double *partial) {
  dreduce_block<force_src, CH>(src, n, B, blockIdx.x, partial);
}

__global__ __launch_bounds__(256) void k_force_cfl_final(const double *partial, int B, double *out, double *cfl,
                                                         double *dt1d, const double *vmoy, const double *newdt,
                                                         const double *invRmed, double coef, double DeltaT,
                                                         int nr, double seq) {
  __shared__ double p[2 * FG_RED_BLOCKS];
  force_cfl_final(partial, B, out, cfl, dt1d, vmoy, newdt, invRmed, coef, DeltaT, nr, seq, p);
}

// results of the reductions for the next substep: CFL count and the disk
// force on the star; copied to pinned memory behind an event
// ring_partials: the star force comes from k_fin's per-ring compensated
// partials (FP32 mode) instead of the DeviceReduce-order sum
static fg_status fused_reductions(fg_ctx *c, double DeltaT, bool ring_partials = false) {
  const double seq = (double)++c->fu_seq;
  c->fu_seq_last = 1;
  if (ring_partials) {
    prof_mark(c);  // (no separate force phase)
    hipLaunchKernelGGL(k_force_rings_cfl, dim3(1), dim3(256), 0, c->stream, c->fu_fpart, c->nr, c->d_fused,
                       c->d_fused + FG_FUSED_CFL_SLOT, c->DT1D, c->Vmoy, c->newDT, c->t.invRmed, c->t.dt1d_coef, DeltaT,
                       seq);
    FG_LAUNCH_CHECK();
  } else {
    // DeviceReduce partials, then one kernel for their final sum and the
    // CFL count
    const long n = (long)c->nr * c->ns;
    const int B = red_blocks(n);
    force_src src{c->f[FG_DENS], c->t.Surf, c->t.Rmed, c->t.cos_cell, c->t.sin_cell, 0.0, 0.0, 0.0, c->ns};
    hipLaunchKernelGGL((k_dreduce_partial_f<RED_CH>), dim3(B), dim3(RED_TPB), 0, c->stream, src, n, B,
                       c->red_partial);
    FG_LAUNCH_CHECK();
    prof_mark(c);
    hipLaunchKernelGGL(k_force_cfl_final, dim3(1), dim3(256), 0, c->stream, c->red_partial, B, c->d_fused,
                       c->d_fused + FG_FUSED_CFL_SLOT, c->DT1D, c->Vmoy, c->newDT, c->t.invRmed, c->t.dt1d_coef,
                       DeltaT, c->nr, seq);
    FG_LAUNCH_CHECK();
  }
  prof_mark(c);
  return FG_OK;
}

// CFL ring pass over the current state (first substep, or after the
// per-routine path touched the state)
__global__ void k_cfl_ring(double *vmoy, double *newdt, const double *vr, const double *vt, const double *cs,
                           const double *dens, const double *energy, int *crash, fg_dev_tables t, int nr, int ns,
                           int fast);

// ring means of a field (any summation order: the FP32 storage's Sigma
// background of the first substep after a state was set)
__global__ __launch_bounds__(256) void k_ring_bg(double *out, const double *f, int ns) {
  __shared__ double w[4];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < ns; j += 256) s += f[(long)i * ns + j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[i] = (w[0] + w[1] + w[2] + w[3]) / (double)ns;
}

extern "C" fg_status fg_fused_prepare(fg_ctx *c, double DeltaT) {
  if (!c) return FG_E_INVALID;
  if (fg_status s = fused_supported(c)) return s;
  if (fg_status s = fused_alloc(c)) return s;
  if (c->p.precision == FG_PREC_FP32) {
    hipLaunchKernelGGL(k_ring_bg, dim3(c->nr), dim3(256), 0, c->stream, c->fu_sbar[c->fu_sbar_cur],
                       (const double *)c->f[FG_DENS], c->ns);
    FG_LAUNCH_CHECK();
  }
  size_t shm = (size_t)(c->ns + 256) * sizeof(double);
  hipLaunchKernelGGL(k_cfl_ring, dim3(c->nr), dim3(256), shm, c->stream, c->Vmoy, c->newDT, c->f[FG_VRAD],
                     c->f[FG_VTHETA], c->f[FG_SOUNDSPEED], c->f[FG_DENS],
                     c->p.adiabatic ? (const double *)c->f[FG_ENERGY] : (const double *)nullptr,
                     (int *)(c->d_fused + FG_FUSED_CRASH_SLOT), c->t, c->nr, c->ns, c->p.fast_transport);
  FG_LAUNCH_CHECK();
  return fused_reductions(c, DeltaT);
}

// Results of the last fused reductions, checked: FG_E_CRASH when a pass
// flagged a negative or non-finite density, or when the CFL count or the
// force shows a time step that went to 0, infinity or NaN (the flag is
// cleared once reported; the next pass over a bad state sets it again).
fg_status fg_fused_check(fg_ctx *c, double cfl, const double *force2) {
  volatile int *flag = (volatile int *)(c->h_fused + FG_FUSED_CRASH_SLOT);
  if (*flag) {
    *flag = 0;
    fg_set_error("crash: negative or non-finite density or energy");
    return FG_E_CRASH;
  }
  return fg_check_substep_results(cfl, force2);
}

fg_status fg_fused_planet_force_async(fg_ctx *c, int k, double x, double y, double rs) {
  if (!c->d_fused || k < 0 || k >= FG_MAX_PLANETS) {
    fg_set_error("fused planet force: no result block or planet %d", k);
    return FG_E_INVALID;
  }
  c->fu_seq_last = 0;  // the results now wait for this reduction too
  return fg_disk_force_async(c, c->f[FG_DENS], x, y, rs, c->d_fused + FG_FUSED_PF_SLOT + 2 * k);
}

void fg_fused_planet_force(const fg_ctx *c, int k, double out2[2]) {
  out2[0] = c->h_fused[FG_FUSED_PF_SLOT + 2 * k];
  out2[1] = c->h_fused[FG_FUSED_PF_SLOT + 2 * k + 1];
}

extern "C" fg_status fg_fused_results(fg_ctx *c, int *ncfl, double *force2, double *lost) {
  if (!c) return FG_E_INVALID;
  if (!c->h_fused) {
    fg_set_error("fg_fused_results: no fused reductions were launched (fg_fused_prepare or fg_step_fused first)");
    return FG_E_INVALID;
  }
  if (c->fu_seq_last && !c->prof_on) {
    // only the final kernel is awaited: watch the slot it writes last
    // (quicker than the stream's completion signal), asking the stream
    // every 1024 polls so that a failed launch still ends the wait
    volatile double *slot = c->h_fused + FG_FUSED_SEQ_SLOT;
    const double want = (double)c->fu_seq;
    for (unsigned n = 1; *slot != want; n++) {
      if ((n & 1023) == 0) {
        const hipError_t e = hipStreamQuery(c->stream);
        if (e == hipErrorNotReady) continue;
        FG_HIP(e);
        if (*slot != want) {
          fg_set_error("fg_fused_results: the final reduction finished without reporting");
          return FG_E_HIP;
        }
        break;
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  } else {
    FG_HIP(hipStreamSynchronize(c->stream));
    prof_fold(c);
  }
  const double cfl = c->h_fused[FG_FUSED_CFL_SLOT];
  const double f[2] = {c->h_fused[0], c->h_fused[1]};
  if (fg_status s = fg_fused_check(c, cfl, f)) return s;
  if (ncfl) *ncfl = (int)cfl;
  if (force2) {
    force2[0] = f[0];
    force2[1] = f[1];
  }
  if (lost) *lost = 0.0;  // evanescent edges: no outflow is recorded
  return FG_OK;
}

// FP32 bulk mode: the rings stored as float pairs, [*b0, *b1): those whose
// centre lies within fp32_pair_hill Hill radii (radially, 1 by default;
// fg_set_fp32_pairs) of a planet, where the circumplanetary material stays
// and the storage rounding would add up (empty without planets or at 0)
static void fp32_pair_rings(const fg_ctx *c, const fg_planet_pot *pl, int nb, int *b0, int *b1) {
  int lo = c->nr, hi = 0;
  for (int k = 0; c->fp32_pair_hill > 0.0 && k < nb; k++) {
    const double rp = sqrt(pl[k].x * pl[k].x + pl[k].y * pl[k].y);
    if (!(pl[k].mass > 0.0) || !(rp > 0.0)) continue;
    const double w = c->fp32_pair_hill * rp * cbrt(pl[k].mass / 3.0);
    for (int i = 0; i < c->nr; i++) {
      const double r = c->h_Rmed[i];
      if (r >= rp - w && r <= rp + w) {
        lo = std::min(lo, i);
        hi = std::max(hi, i + 1);
      }
    }
  }
  if (lo >= hi) lo = hi = 0;
  *b0 = lo;
  *b1 = hi;
}

extern "C" fg_status fg_step_fused(fg_ctx *c, double dt, double omega_frame, double domega,
                                   const fg_planet_pot *planets, int nb, double ix, double iy,
                                   double DeltaT_next, int flags) {
  if (!c) return FG_E_INVALID;
  if (fg_status s = fused_supported(c)) return s;
  if (fg_status s = fused_alloc(c)) return s;
  if (nb < 0 || nb > FG_MAX_PLANETS || (nb > 0 && !planets)) {
    fg_set_error("fg_step_fused: planet count %d", nb);
    return FG_E_INVALID;
  }
  const int want_temperature = flags & FG_STEP_TEMPERATURE;
  const int nr = c->nr, ns = c->ns;
  const bool adi = c->p.adiabatic, sg = c->p.self_gravity;
  fused_bufs b = bufs_of(c);
  // this substep's viscosity: refreshed at the close of the last DT, it
  // replaces nu after the first damping pass
  const double *nu = c->nu_pending ? c->nu_next : c->nu;
  if (sg)
    if (fg_status s = fg_sg_fused(c, dt, b.kick_r, b.kick_t)) return s;
  // k_src
  src_args sa;
  sa.dens = c->f[FG_DENS];
  sa.vr = c->f[FG_VRAD];
  sa.vt = c->f[FG_VTHETA];
  sa.odens = b.Bd;
  sa.ovr = b.Bvr;
  sa.ovt = b.Bvt;
  sa.opress = (want_temperature && !adi) ? b.P : nullptr;
  sa.nu = nu;
  sa.nu1 = c->nu;
  sa.energy = adi ? c->f[FG_ENERGY] : nullptr;
  sa.oeint = b.Beint;
  sa.gamma = c->p.adiabatic_index;
  sa.kick_r = b.kick_r;
  sa.kick_t = b.kick_t;
  sa.sg1 = sg ? c->sgacc_prev : nullptr;
  sa.sg2 = sg ? c->sgacc_mean : nullptr;
  sa.t = c->t;
  sa.P.nb = nb;
  for (int k = 0; k < nb; k++) {
    sa.P.x[k] = planets[k].x;
    sa.P.y[k] = planets[k].y;
    sa.P.m[k] = planets[k].mass;
    sa.P.smooth[k] = planets[k].smooth;
    sa.P.invd3[k] = planets[k].inv_d3;
  }
  sa.P.ix = ix;
  sa.P.iy = iy;
  sa.dt = dt;
  sa.omega_frame = omega_frame;
  sa.domega = domega;
  sa.nr = nr;
  sa.ns = ns;
  sa.indirect = c->p.indirect_term;
  sa.evanescent = c->p.evanescent;
  sa.corotating = (domega != 0.0) ? 1 : 0;
  if (c->prof_on) {
    (void)hipStreamSynchronize(c->stream);
    prof_fold(c);
    c->prof_nev = 0;
    c->prof_pending = 1;
  }
  const bool fp32 = c->p.precision == FG_PREC_FP32;
  int lob0 = 0, lob1 = 0;
  if (fp32) fp32_pair_rings(c, planets, nb, &lob0, &lob1);
  prof_mark(c);
  if (fp32) {
    src32_args s3;
    s3.lob0 = lob0;
    s3.lob1 = lob1;
    s3.dens = sa.dens;
    s3.vr = sa.vr;
    s3.vt = sa.vt;
    s3.vref = c->Vmoy;  // ring means of vtheta from the previous substep
    s3.sbar = c->fu_sbar[c->fu_sbar_cur];
    s3.odens = b.fBd();
    s3.ovr = b.fBvr();
    s3.ovt = b.fBvt();
    s3.opress = sa.opress;
    s3.nu = sa.nu;
    s3.t = sa.t;
    s3.P = sa.P;
    s3.dt = dt;
    s3.omega_frame = omega_frame;
    s3.domega = domega;
    s3.nr = nr;
    s3.ns = ns;
    s3.indirect = sa.indirect;
    s3.evanescent = sa.evanescent;
    s3.corotating = sa.corotating;
    const int th = c->tile_src_th;
    FG_HIP(hipFuncSetAttribute((const void *)k_src32<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)src32_lds_bytes(32)));
    const dim3 g32((ns + TW - 1) / TW, (nr + 1 + th - 1) / th);
    if (th == 8)
      hipLaunchKernelGGL(k_src32<8>, g32, dim3(SRC32_TPB), src32_lds_bytes(8), c->stream, s3);
    else if (th == 32)
      hipLaunchKernelGGL(k_src32<32>, g32, dim3(SRC32_TPB), src32_lds_bytes(32), c->stream, s3);
    else
      hipLaunchKernelGGL(k_src32<16>, g32, dim3(SRC32_TPB), src32_lds_bytes(16), c->stream, s3);
  } else {
    const dim3 gs((ns + TW - 1) / TW, (nr + 1 + TH - 1) / TH);
    const int mode = (adi ? 1 : 0) | (sg ? 2 : 0);
    if (mode == 0)
      hipLaunchKernelGGL((k_src<TH, TW, SRC64_TPB, 0>), gs, dim3(SRC64_TPB), 0, c->stream, sa);
    else if (mode == 1)
      hipLaunchKernelGGL((k_src<TH, TW, SRC64_TPB, 1>), gs, dim3(SRC64_TPB), 0, c->stream, sa);
    else if (mode == 2)
      hipLaunchKernelGGL((k_src<TH, TW, SRC64_TPB, 2>), gs, dim3(SRC64_TPB), 0, c->stream, sa);
    else
      hipLaunchKernelGGL((k_src<TH, TW, SRC64_TPB, 3>), gs, dim3(SRC64_TPB), 0, c->stream, sa);
  }
  FG_LAUNCH_CHECK();
  if (adi) {  // Substep3 (counted with k_src by fg_profile)
    src3_args s3;
    s3.dens = b.Bd;
    s3.vr = b.Bvr;
    s3.vt = b.Bvt;
    s3.eint = b.Beint;
    s3.energy = c->f[FG_ENERGY];
    s3.nu = nu;
    s3.t = c->t;
    s3.dt = dt;
    s3.gamma = c->p.adiabatic_index;
    s3.lr01 = c->lr01;
    s3.lr12 = c->lr12;
    s3.nr = nr;
    s3.ns = ns;
    s3.cooling = c->p.cooling;
    hipLaunchKernelGGL(k_src3, dim3((ns + 63) / 64, (nr + 3) / 4), dim3(256), 0, c->stream, s3);
    FG_LAUNCH_CHECK();
  }
  // k_trad
  trad_args ta;
  ta.dens = b.Bd;
  ta.vr = b.Bvr;
  ta.vt = b.Bvt;
  ta.energy = adi ? c->f[FG_ENERGY] : nullptr;
  ta.out[0] = b.Rrmp;
  ta.out[1] = b.Rrmm;
  ta.out[2] = b.Rtmp;
  ta.out[3] = b.Rtmm;
  ta.out[4] = adi ? b.Re : b.Rd;
  ta.out[5] = b.Rd;
  ta.fdens = b.fBd();
  ta.fvr = b.fBvr();
  ta.fvt = b.fBvt();
  for (int q = 0; q < NQ; q++) ta.fout[q] = b.fR(q);
  ta.sbar = c->fu_sbar[c->fu_sbar_cur];
  ta.vref = c->Vmoy;
  ta.t = c->t;
  ta.dt = dt;
  ta.omega_frame = omega_frame;
  ta.nr = nr;
  ta.ns = ns;
  ta.vmed = c->VMed;
  ta.nstrips = (ns + TRAD_TPB - 1) / TRAD_TPB;
  ta.band = c->tile_trad_band;
  ta.lob0 = lob0;
  ta.lob1 = lob1;
  prof_mark(c);
  const dim3 gtrad(nr + ta.nstrips * ((nr + ta.band - 1) / ta.band));
  if (fp32)
    hipLaunchKernelGGL(k_trad32, gtrad, dim3(TRAD_TPB), 0, c->stream, ta);
  else if (adi)
    hipLaunchKernelGGL(k_trad<6>, gtrad, dim3(TRAD_TPB), 0, c->stream, ta);
  else
    hipLaunchKernelGGL(k_trad<5>, gtrad, dim3(TRAD_TPB), 0, c->stream, ta);
  FG_LAUNCH_CHECK();
  // k_taz
  taz_args za;
  za.vt = b.Bvt;
  za.vmed = c->VMed;
  za.q[0] = b.Rrmp;
  za.q[1] = b.Rrmm;
  za.q[2] = b.Rtmp;
  za.q[3] = b.Rtmm;
  za.q[4] = adi ? b.Re : b.Rd;
  za.q[5] = b.Rd;
  za.xd = b.Xd;
  za.xvt = b.Xvt;
  za.xrmp = b.Xrmp;
  za.xrmm = b.Xrmm;
  za.xe = b.Xe;
  za.fvt = b.fBvt();
  for (int q = 0; q < NQ; q++) za.fq[q] = b.fR(q);
  za.fxd = b.fXd();
  za.fxvt = b.fXvt();
  za.fxrmp = b.fXrmp();
  za.fxrmm = b.fXrmm();
  za.sbar = c->fu_sbar[c->fu_sbar_cur];
  za.vref = c->Vmoy;
  za.nshift = c->Nshift;
  za.t = c->t;
  za.dt = dt;
  za.omega_frame = omega_frame;
  za.nr = nr;
  za.ns = ns;
  za.fast = c->p.fast_transport;
  za.lob0 = lob0;
  za.lob1 = lob1;
  prof_mark(c);
  if (!fp32) {
    // FP64, five quantities or six with the energy: at 1024 x 3072, 17
    // chunks a ring (6 % halo cells) at three waves per SIMD run 65 us
    // against 72 us for 10 chunks of five cells per lane at two waves (the
    // ring as one 256-thread workgroup took 143 us)
    const dim3 gw((ns + TAZW_P64 - 1) / TAZW_P64, nr);
    if (adi)
      hipLaunchKernelGGL((k_tazw<TAZW_CL64, false, 6>), gw, dim3(64), 0, c->stream, za);
    else
      hipLaunchKernelGGL((k_tazw<TAZW_CL64, false>), gw, dim3(64), 0, c->stream, za);
  } else {
    // FP32: three cells per lane at five waves per SIMD, 40.7 -> 38.4 us at
    // 1024 x 3072 against four cells per lane at four (113 us as one
    // 256-thread workgroup per ring)
    const dim3 gw((ns + TAZW_P32S - 1) / TAZW_P32S, nr);
    hipLaunchKernelGGL(k_tazw32, gw, dim3(64), 0, c->stream, za);
  }
  FG_LAUNCH_CHECK();
  // k_fin
  fin_args fa;
  fa.xd = b.Xd;
  fa.xvt = b.Xvt;
  fa.xrmp = b.Xrmp;
  fa.xrmm = b.Xrmm;
  fa.press = b.P;
  fa.dens = c->f[FG_DENS];
  fa.vr = c->f[FG_VRAD];
  fa.vt = c->f[FG_VTHETA];
  fa.temp = want_temperature ? c->f[FG_TEMPERATURE] : nullptr;
  fa.vmoy = c->Vmoy;
  fa.newdt = c->newDT;
  fa.nu = c->nu;
  fa.t = c->t;
  fa.dt = dt;
  fa.omega_frame = omega_frame;
  fa.nr = nr;
  fa.ns = ns;
  fa.evanescent = c->p.evanescent;
  fa.fast = c->p.fast_transport;
  fa.fpart = fp32 ? c->fu_fpart : nullptr;
  fa.crash = (int *)(c->d_fused + FG_FUSED_CRASH_SLOT);
  fa.nu = nu;
  fa.xe = b.Xe;
  fa.energy = adi ? c->f[FG_ENERGY] : nullptr;
  fa.cs = c->f[FG_SOUNDSPEED];
  fa.gamma = c->p.adiabatic_index;
  fa.adiabatic = adi ? 1 : 0;
  fa.cs_refresh = (adi && (flags & FG_STEP_CLOSES_DT)) ? 1 : 0;
  fa.sgacc = sg ? c->sgacc_mean : nullptr;
  fa.fxd = b.fXd();
  fa.fxvt = b.fXvt();
  fa.fxrmp = b.fXrmp();
  fa.fxrmm = b.fXrmm();
  fa.sbar = c->fu_sbar[c->fu_sbar_cur];
  fa.sbar_next = c->fu_sbar[c->fu_sbar_cur ^ 1];
  fa.lob0 = lob0;
  fa.lob1 = lob1;
  prof_mark(c);
  const dim3 gfin(FG_NXCD * ((nr + FG_NXCD - 1) / FG_NXCD));
  if (adi)
    hipLaunchKernelGGL(k_fin<true>, gfin, dim3(256), (size_t)ns * sizeof(double), c->stream, fa);
  else if (fp32)
    hipLaunchKernelGGL((k_fin<false, true>), gfin, dim3(256), (size_t)ns * sizeof(double), c->stream, fa);
  else
    hipLaunchKernelGGL(k_fin<false>, gfin, dim3(256), (size_t)ns * sizeof(double), c->stream, fa);
  FG_LAUNCH_CHECK();
  if (fp32) c->fu_sbar_cur ^= 1;  // k_fin wrote the next substep's Sigma background
  prof_mark(c);
  if (c->nu_pending) {  // the refreshed viscosity is now the current one
    std::swap(c->nu, c->nu_next);
    c->nu_pending = 0;
  }
  if (fg_status s = fused_reductions(c, DeltaT_next, fp32)) return s;
  if (fa.cs_refresh && c->p.viscosity_alpha) {
    // the next DT's alpha viscosity from the ring means of the new sound
    // speed (FViscosity via Make1Dprofile, as fg_ring_mean_cs computes it);
    // the next substep's first damping pass still uses this one
    if (fg_status s = fg_viscosity_into(c, c->nu_next)) return s;
    c->nu_pending = 1;
  }
  return FG_OK;
}

// ================================================== batched ensemble ====
// Many independent disks (uneven grids, FP32 bulk mode) advanced one
// substep together: each fused phase is ONE launch over a work list of
// (disk, tile) items whose workgroups stride through the list, so small
// and large disks share the machine.  Per substep the host writes every
// disk's arguments (dt, planets, frame, active flag) into one pinned
// block; a hipGraph captured once copies it to the device, runs the
// phases and the reductions, and copies back each disk's CFL count and
// star force.  The per-disk arithmetic is the single-disk FP32 code, so
// a disk gives the same bits batched or alone.

struct batch_disk {
  src32_args src;
  trad_args trad;
  taz_args taz;
  fin_args fin;
  force_src force;
  long ncell;      // NRAD * NSEC
  int B;           // DeviceReduce blocks of the force sum
  double *partial; // block sums (the context's reduction scratch)
  double *DT1D, *Vmoy, *newDT;
  const double *invRmed;
  double coef, DeltaT;  // CFL coefficient and the next substep's DeltaT
  int nr, active;
};

struct batch_result {
  double fx, fy;
  int cfl, pad;
};

// Each phase launch covers its whole work list, one workgroup per item (the
// kernels keep a grid-stride loop, so any grid is correct).  Measured on
// 64 disks: 1.83 ms per batch substep with 4096 looping workgroups per
// phase, 1.61 ms with one item per workgroup.
#define BATCH_CAP_BLOCKS (1 << 20)

// work items are the same for the whole workgroup: keep them (and every
// address derived from them) in scalar registers
__device__ __forceinline__ int4 uniform4(int4 w) {
  return make_int4(__builtin_amdgcn_readfirstlane(w.x), __builtin_amdgcn_readfirstlane(w.y),
                   __builtin_amdgcn_readfirstlane(w.z), __builtin_amdgcn_readfirstlane(w.w));
}
__device__ __forceinline__ int2 uniform2(int2 w) {
  return make_int2(__builtin_amdgcn_readfirstlane(w.x), __builtin_amdgcn_readfirstlane(w.y));
}

// one work item per workgroup: without a loop's carried scalars the tile's
// uniform arguments fit the SGPRs instead of spilling into VGPR lanes
// (253 -> 112 VGPRs)
template <int ROWS>
__global__ __launch_bounds__(SRC32_TPB) void k_batch_src(const batch_disk *__restrict__ D, const int4 *__restrict__ items, int n) {
  extern __shared__ float lds_bsrc[];
  if ((int)blockIdx.x >= n) return;
  const int4 w = uniform4(items[blockIdx.x]);
  if (D[w.x].active) src32_tile<ROWS>(D[w.x].src, w.y, w.z, lds_bsrc);
}

__global__ __launch_bounds__(TRAD_TPB) void k_batch_trad(const batch_disk *__restrict__ D, const int2 *__restrict__ items, int n) {
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = uniform2(items[it]);
    if (D[w.x].active) trad32_block(D[w.x].trad, w.y);
    __syncthreads();
  }
}

// one chunk of one ring per item: y = ring * chunks + chunk
__global__ __launch_bounds__(64) void k_batch_taz(const batch_disk *__restrict__ D, const int2 *__restrict__ items, int n) {
  __shared__ tazw_lds<4, 5> sh;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = uniform2(items[it]);
    if (D[w.x].active) {
      const int nch = (D[w.x].taz.ns + TAZW_P32 - 1) / TAZW_P32;
      const int i = w.y / nch;
      if (in_lob(i, D[w.x].taz.lob0, D[w.x].taz.lob1))
        tazw_body<4, true, 5, true>(D[w.x].taz, i, w.y % nch, sh);
      else
        tazw_body<4, true, 5, false>(D[w.x].taz, i, w.y % nch, sh);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_fin(const batch_disk *__restrict__ D, const int2 *__restrict__ items, int n) {
  extern __shared__ double lds_bfin[];
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = uniform2(items[it]);
    if (D[w.x].active) {
      const fin_args &fa = D[w.x].fin;
      if (in_lob(w.y, fa.lob0 - 1, fa.lob1 + 1))
        fin_ring<1, false, true, true>(fa, w.y, lds_bfin);
      else
        fin_ring<1, false, true, false>(fa, w.y, lds_bfin);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(RED_TPB) void k_batch_force(const batch_disk *__restrict__ D, const int2 *__restrict__ items, int n) {
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = uniform2(items[it]);
    const batch_disk &d = D[w.x];
    if (d.active) dreduce_block(d.force, d.ncell, d.B, w.y, d.partial);
    __syncthreads();
  }
}

// per disk: the force's final sum and the CFL count of the next substep
__global__ __launch_bounds__(256) void k_batch_close(const batch_disk *__restrict__ D, batch_result *__restrict__ res, int ndisk) {
  __shared__ double out[2];
  for (int k = blockIdx.x; k < ndisk; k += gridDim.x) {
    const batch_disk &d = D[k];
    if (d.active) {
      force_rings_final(d.fin.fpart, d.nr, out);
      __syncthreads();
      cfl_final(&res[k].cfl, d.DT1D, d.Vmoy, d.newDT, d.invRmed, d.coef, d.DeltaT, d.nr);
      if (threadIdx.x == 0) {
        res[k].fx = out[0];
        res[k].fy = out[1];
      }
    }
    __syncthreads();
  }
}

struct fg_batch_dev {
  int n;
  std::vector<fg_ctx *> ctx;
  hipStream_t stream;
  batch_disk *h_disk, *d_disk;  // pinned host block and its device copy
  batch_result *h_res, *d_res;
  int4 *it_src;
  int2 *it_trad, *it_taz, *it_fin, *it_force;
  int n_src, n_trad, n_taz, n_fin, n_force;
  int n_src_rows[3];  // source items of tile height 8, 16, 32 (in that order in it_src)
  int max_ns;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int ready;  // items built and graph instantiated
};

static void batch_free(fg_batch_dev *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void *dev[] = {b->d_disk, b->d_res, b->it_src, b->it_trad, b->it_taz, b->it_fin, b->it_force};
  for (void *q : dev)
    if (q) (void)hipFree(q);
  if (b->h_disk) (void)hipHostFree(b->h_disk);
  if (b->h_res) (void)hipHostFree(b->h_res);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

void fg_batch_dev_destroy(fg_batch_dev *b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
}

fg_status fg_batch_dev_create(fg_ctx *const *ctx, int n, fg_batch_dev **out) {
  *out = nullptr;
  if (n <= 0) {
    fg_set_error("batch: no disks");
    return FG_E_INVALID;
  }
  for (int k = 0; k < n; k++) {
    if (!ctx[k] || ctx[k]->p.precision != FG_PREC_FP32) {
      fg_set_error("batch: disk %d is not an FP32 bulk-mode context", k);
      return FG_E_INVALID;
    }
    if (fg_status s = fused_supported(ctx[k])) return s;
    if (fg_status s = fused_alloc(ctx[k])) return s;
    if (ctx[k]->device != ctx[0]->device) {
      fg_set_error("batch: all disks must live on one device");
      return FG_E_INVALID;
    }
  }
  fg_batch_dev *b = new fg_batch_dev();
  b->n = n;
  b->ctx.assign(ctx, ctx + n);
  hipError_t e = hipStreamCreate(&b->stream);
  if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_disk, n * sizeof(batch_disk), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_res, n * sizeof(batch_result), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void **)&b->d_disk, n * sizeof(batch_disk));
  if (e == hipSuccess) e = hipMalloc((void **)&b->d_res, n * sizeof(batch_result));
  if (e != hipSuccess) {
    fg_set_error("batch: %s", hipGetErrorString(e));
    batch_free(b);
    return FG_E_HIP;
  }
  std::memset((void *)b->h_disk, 0, n * sizeof(batch_disk));
  std::memset((void *)b->h_res, 0, n * sizeof(batch_result));
  *out = b;
  return FG_OK;
}

template <class T>
static fg_status upload_items(const std::vector<T> &v, T **dev) {
  FG_HIP(hipMalloc((void **)dev, (v.empty() ? 1 : v.size()) * sizeof(T)));
  if (!v.empty()) FG_HIP(hipMemcpy(*dev, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return FG_OK;
}

// work lists (largest disks first) and the captured graph of one substep
static fg_status batch_build(fg_batch_dev *b) {
  std::vector<int> order(b->n);
  for (int k = 0; k < b->n; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->ctx[x]->S > b->ctx[y]->S; });
  std::vector<int4> src_rows[3];  // by tile height 8, 16, 32
  std::vector<int2> trad, taz, fin, force;
  b->max_ns = 1;
  for (int k : order) {
    const fg_ctx *c = b->ctx[k];
    const int nr = c->nr, ns = c->ns, th = c->tile_src_th, band = c->tile_trad_band;
    if ((th != 8 && th != 16 && th != 32) || band < 1) {
      fg_set_error("batch: disk %d has tile shape th=%d band=%d", k, th, band);
      return FG_E_INVALID;
    }
    b->max_ns = std::max(b->max_ns, ns);
    std::vector<int4> &src = src_rows[th == 8 ? 0 : th == 16 ? 1 : 2];
    for (int by = 0; by < (nr + 1 + th - 1) / th; by++)
      for (int bx = 0; bx < (ns + TW - 1) / TW; bx++) src.push_back(make_int4(k, bx, by, th));
    const int nstrips = (ns + TRAD_TPB - 1) / TRAD_TPB;
    const int nblk = nr + nstrips * ((nr + band - 1) / band);
    for (int q = 0; q < nblk; q++) trad.push_back(make_int2(k, q));
    const int nch = (ns + TAZW_P32 - 1) / TAZW_P32;
    for (int i = 0; i < nr; i++) {
      for (int ch = 0; ch < nch; ch++) taz.push_back(make_int2(k, i * nch + ch));
      fin.push_back(make_int2(k, i));
    }
    const int B = red_blocks((long)nr * ns);
    for (int q = 0; q < B; q++) force.push_back(make_int2(k, q));
  }
  std::vector<int4> src;
  for (int h = 0; h < 3; h++) {
    b->n_src_rows[h] = (int)src_rows[h].size();
    src.insert(src.end(), src_rows[h].begin(), src_rows[h].end());
  }
  b->n_src = (int)src.size();
  b->n_trad = (int)trad.size();
  b->n_taz = (int)taz.size();
  b->n_fin = (int)fin.size();
  b->n_force = (int)force.size();
  if (fg_status s = upload_items(src, &b->it_src)) return s;
  if (fg_status s = upload_items(trad, &b->it_trad)) return s;
  if (fg_status s = upload_items(taz, &b->it_taz)) return s;
  if (fg_status s = upload_items(fin, &b->it_fin)) return s;
  if (fg_status s = upload_items(force, &b->it_force)) return s;
  FG_HIP(hipFuncSetAttribute((const void *)k_batch_src<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)src32_lds_bytes(32)));
  // capture: arguments in, five phases, two reductions, results out
  hipStream_t st = b->stream;
  FG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  auto grid = [](int n) { return dim3(n < BATCH_CAP_BLOCKS ? (n > 0 ? n : 1) : BATCH_CAP_BLOCKS); };
  (void)hipMemcpyAsync(b->d_disk, b->h_disk, b->n * sizeof(batch_disk), hipMemcpyHostToDevice, st);
  // one source launch per tile height present (the work list holds them
  // in order 8, 16, 32)
  {
    const int4 *it = b->it_src;
    const int n8 = b->n_src_rows[0], n16 = b->n_src_rows[1], n32 = b->n_src_rows[2];
    if (n8) hipLaunchKernelGGL(k_batch_src<8>, dim3(n8), dim3(SRC32_TPB), src32_lds_bytes(8), st, b->d_disk, it, n8);
    if (n16)
      hipLaunchKernelGGL(k_batch_src<16>, dim3(n16), dim3(SRC32_TPB), src32_lds_bytes(16), st, b->d_disk, it + n8, n16);
    if (n32)
      hipLaunchKernelGGL(k_batch_src<32>, dim3(n32), dim3(SRC32_TPB), src32_lds_bytes(32), st, b->d_disk,
                         it + n8 + n16, n32);
  }
  hipLaunchKernelGGL(k_batch_trad, grid(b->n_trad), dim3(TRAD_TPB), 0, st, b->d_disk, b->it_trad, b->n_trad);
  hipLaunchKernelGGL(k_batch_taz, grid(b->n_taz), dim3(64), 0, st, b->d_disk, b->it_taz, b->n_taz);
  hipLaunchKernelGGL(k_batch_fin, grid(b->n_fin), dim3(256), (size_t)b->max_ns * sizeof(double), st, b->d_disk,
                     b->it_fin, b->n_fin);
  hipLaunchKernelGGL(k_batch_close, grid(b->n), dim3(256), 0, st, b->d_disk, b->d_res, b->n);
  (void)hipMemcpyAsync(b->h_res, b->d_res, b->n * sizeof(batch_result), hipMemcpyDeviceToHost, st);
  hipError_t ce = hipStreamEndCapture(st, &b->graph);
  if (ce != hipSuccess) {
    fg_set_error("batch graph capture: %s", hipGetErrorString(ce));
    return FG_E_HIP;
  }
  FG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->ready = 1;
  return FG_OK;
}

// arguments of disk k for its next substep (FP32 fused kernels); active=0
// leaves the disk untouched this substep
fg_status fg_batch_dev_set(fg_batch_dev *b, int k, int active, double dt, double omega_frame, double domega,
                           const fg_planet_pot *planets, int nb, double ix, double iy, double DeltaT_next) {
  if (!b || k < 0 || k >= b->n) return FG_E_INVALID;
  batch_disk &d = b->h_disk[k];
  d.active = active;
  if (!active) return FG_OK;
  if (nb < 0 || nb > FG_MAX_PLANETS || (nb > 0 && !planets)) {
    fg_set_error("batch: planet count %d", nb);
    return FG_E_INVALID;
  }
  fg_ctx *c = b->ctx[k];
  const int nr = c->nr, ns = c->ns;
  fused_bufs fb = bufs_of(c);
  int lob0 = 0, lob1 = 0;
  fp32_pair_rings(c, planets, nb, &lob0, &lob1);
  src32_args &s = d.src;
  s.lob0 = lob0;
  s.lob1 = lob1;
  s.dens = c->f[FG_DENS];
  s.vr = c->f[FG_VRAD];
  s.vt = c->f[FG_VTHETA];
  // FP32 storage backgrounds: this substep reads fu_sbar[cur], its k_fin
  // writes the other one for the next
  const double *sbar = c->fu_sbar[c->fu_sbar_cur];
  double *sbar_next = c->fu_sbar[c->fu_sbar_cur ^ 1];
  c->fu_sbar_cur ^= 1;
  s.vref = c->Vmoy;
  s.sbar = sbar;
  s.odens = fb.fBd();
  s.ovr = fb.fBvr();
  s.ovt = fb.fBvt();
  s.opress = nullptr;
  s.nu = c->nu;
  s.t = c->t;
  s.P.nb = nb;
  for (int q = 0; q < nb; q++) {
    s.P.x[q] = planets[q].x;
    s.P.y[q] = planets[q].y;
    s.P.m[q] = planets[q].mass;
    s.P.smooth[q] = planets[q].smooth;
    s.P.invd3[q] = planets[q].inv_d3;
  }
  s.P.ix = ix;
  s.P.iy = iy;
  s.dt = dt;
  s.omega_frame = omega_frame;
  s.domega = domega;
  s.nr = nr;
  s.ns = ns;
  s.indirect = c->p.indirect_term;
  s.evanescent = c->p.evanescent;
  s.corotating = (domega != 0.0) ? 1 : 0;
  trad_args &ta = d.trad;
  ta.dens = fb.Bd;
  ta.vr = fb.Bvr;
  ta.vt = fb.Bvt;
  ta.out[0] = fb.Rrmp;
  ta.out[1] = fb.Rrmm;
  ta.out[2] = fb.Rtmp;
  ta.out[3] = fb.Rtmm;
  ta.out[4] = fb.Rd;
  ta.fdens = fb.fBd();
  ta.fvr = fb.fBvr();
  ta.fvt = fb.fBvt();
  for (int q = 0; q < NQ; q++) ta.fout[q] = fb.fR(q);
  ta.sbar = sbar;
  ta.vref = c->Vmoy;
  ta.vmed = c->VMed;
  ta.t = c->t;
  ta.dt = dt;
  ta.omega_frame = omega_frame;
  ta.nr = nr;
  ta.ns = ns;
  ta.nstrips = (ns + TRAD_TPB - 1) / TRAD_TPB;
  ta.band = c->tile_trad_band;
  ta.lob0 = lob0;
  ta.lob1 = lob1;
  taz_args &za = d.taz;
  za.vt = fb.Bvt;
  za.vmed = c->VMed;
  za.q[0] = fb.Rrmp;
  za.q[1] = fb.Rrmm;
  za.q[2] = fb.Rtmp;
  za.q[3] = fb.Rtmm;
  za.q[4] = fb.Rd;
  za.xd = fb.Xd;
  za.xvt = fb.Xvt;
  za.xrmp = fb.Xrmp;
  za.xrmm = fb.Xrmm;
  za.fvt = fb.fBvt();
  for (int q = 0; q < NQ; q++) za.fq[q] = fb.fR(q);
  za.fxd = fb.fXd();
  za.fxvt = fb.fXvt();
  za.fxrmp = fb.fXrmp();
  za.fxrmm = fb.fXrmm();
  za.sbar = sbar;
  za.vref = c->Vmoy;
  za.nshift = c->Nshift;
  za.t = c->t;
  za.dt = dt;
  za.omega_frame = omega_frame;
  za.nr = nr;
  za.ns = ns;
  za.fast = c->p.fast_transport;
  za.lob0 = lob0;
  za.lob1 = lob1;
  fin_args &fa = d.fin;
  fa.xd = fb.Xd;
  fa.xvt = fb.Xvt;
  fa.xrmp = fb.Xrmp;
  fa.xrmm = fb.Xrmm;
  fa.press = fb.P;
  fa.dens = c->f[FG_DENS];
  fa.vr = c->f[FG_VRAD];
  fa.vt = c->f[FG_VTHETA];
  fa.temp = nullptr;
  fa.vmoy = c->Vmoy;
  fa.newdt = c->newDT;
  fa.nu = c->nu;
  fa.t = c->t;
  fa.dt = dt;
  fa.omega_frame = omega_frame;
  fa.nr = nr;
  fa.ns = ns;
  fa.evanescent = c->p.evanescent;
  fa.fast = c->p.fast_transport;
  fa.fpart = c->fu_fpart;
  fa.crash = (int *)(c->d_fused + FG_FUSED_CRASH_SLOT);
  fa.xe = nullptr;
  fa.energy = fa.cs = nullptr;
  fa.gamma = 0.0;
  fa.adiabatic = fa.cs_refresh = 0;
  fa.sgacc = nullptr;
  fa.fxd = fb.fXd();
  fa.fxvt = fb.fXvt();
  fa.fxrmp = fb.fXrmp();
  fa.fxrmm = fb.fXrmm();
  fa.sbar = sbar;
  fa.sbar_next = sbar_next;
  fa.lob0 = lob0;
  fa.lob1 = lob1;
  d.force = force_src{c->f[FG_DENS], c->t.Surf, c->t.Rmed, c->t.cos_cell, c->t.sin_cell, 0.0, 0.0, 0.0, ns};
  d.ncell = (long)nr * ns;
  d.B = red_blocks(d.ncell);
  d.partial = c->red_partial;
  d.DT1D = c->DT1D;
  d.Vmoy = c->Vmoy;
  d.newDT = c->newDT;
  d.invRmed = c->t.invRmed;
  d.coef = c->t.dt1d_coef;
  d.DeltaT = DeltaT_next;
  d.nr = nr;
  return FG_OK;
}

// run one substep of every active disk (asynchronous)
fg_status fg_batch_dev_launch(fg_batch_dev *b) {
  if (!b->ready)
    if (fg_status s = batch_build(b)) return s;
  FG_HIP(hipGraphLaunch(b->exec, b->stream));
  return FG_OK;
}

// wait for the last launch; results of disk k
fg_status fg_batch_dev_wait(fg_batch_dev *b) {
  FG_HIP(hipStreamSynchronize(b->stream));
  return FG_OK;
}

void fg_batch_dev_result(const fg_batch_dev *b, int k, int *cfl, double *force2) {
  *cfl = b->h_res[k].cfl;
  force2[0] = b->h_res[k].fx;
  force2[1] = b->h_res[k].fy;
}

// after a tile-shape change the work lists and the graph are rebuilt
void fg_batch_dev_invalidate(fg_batch_dev *b) {
  if (!b || !b->ready) return;
  (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (void **q : {(void **)&b->it_src, (void **)&b->it_trad, (void **)&b->it_taz, (void **)&b->it_fin,
                   (void **)&b->it_force}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  b->ready = 0;
}

extern "C" fg_status fg_set_fp32_pairs(fg_ctx *c, double hill_radii) {
  if (!c) return FG_E_INVALID;
  if (!(hill_radii >= 0.0 && hill_radii <= 100.0)) {
    fg_set_error("fg_set_fp32_pairs: %g Hill radii (0..100)", hill_radii);
    return FG_E_INVALID;
  }
  c->fp32_pair_hill = hill_radii;
  return FG_OK;
}

extern "C" fg_status fg_set_tiles(fg_ctx *c, int src_rows, int trad_band) {
  if (!c) return FG_E_INVALID;
  if ((src_rows != 8 && src_rows != 16 && src_rows != 32) || trad_band < 1 || trad_band > 4096) {
    fg_set_error("fg_set_tiles: rows %d (8, 16 or 32), band %d (1..4096)", src_rows, trad_band);
    return FG_E_INVALID;
  }
  c->tile_src_th = src_rows;
  c->tile_trad_band = trad_band;
  return FG_OK;
}

// ------------------------------------------------- shared-divisor check ----
// bitwise comparison of recip64(d).div(n) with the division n / d
__global__ void k_check_recip64(const double *num, const double *den, long n, unsigned long long *bad) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double q = num[t] / den[t];
  const double s = recip64(den[t]).div(num[t]);
  if (__double_as_longlong(q) != __double_as_longlong(s)) atomicAdd(bad, 1ull);
}

extern "C" fg_status fg_check_shared_division(const double *num, const double *den, long n, long *mismatches) {
  if (!num || !den || !mismatches || n < 0) return FG_E_INVALID;
  *mismatches = 0;
  if (n == 0) return FG_OK;
  double *dn = nullptr, *dd = nullptr;
  unsigned long long *dbad = nullptr, hbad = 0;
  hipError_t e = hipMalloc((void **)&dn, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&dd, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&dbad, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemcpy(dn, num, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd, den, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbad, &hbad, sizeof(hbad), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_check_recip64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dn, dd, n, dbad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost);
  if (dn) (void)hipFree(dn);
  if (dd) (void)hipFree(dd);
  if (dbad) (void)hipFree(dbad);
  if (e != hipSuccess) {
    fg_set_error("fg_check_shared_division: %s", hipGetErrorString(e));
    return FG_E_HIP;
  }
  *mismatches = (long)hbad;
  return FG_OK;
}
