// Per-routine path: one HIP kernel (or a few) per reference routine, in
// the reference's order.  This is the drop-in granularity of the C-ABI
// (fg_pressure, fg_substep1, ...); the fused path in fg_fused.hip computes
// the same values with fewer passes.
//
// Launch shape: 1D, 256 threads, one thread per cell l = i*ns + j (row-
// major, sectors contiguous, so a wavefront reads 64 consecutive doubles).
#include "fg_cell.h"
#include "fg_internal.h"
#include "fg_reduce.h"

#define TPB 256

#define CELLS(i0, i1)                                                  \
  long l = (long)blockIdx.x * blockDim.x + threadIdx.x + (long)(i0)*ns; \
  if (l >= (long)(i1)*ns) return;                                      \
  const int i = (int)(l / ns);                                         \
  const int j = (int)(l - (long)i * ns);                               \
  (void)j;

#define JP ((long)i * ns + ((j + 1) % ns))
#define JM ((long)i * ns + ((j - 1 + ns) % ns))

#define LAUNCH(kern, i0, i1, ...)                                                              \
  do {                                                                                         \
    long n_ = (long)((i1) - (i0)) * c->ns;                                                     \
    if (n_ > 0) hipLaunchKernelGGL(kern, dim3(fg_blocks(n_, TPB)), dim3(TPB), 0, c->stream, __VA_ARGS__); \
    FG_LAUNCH_CHECK();                                                                         \
  } while (0)

static inline double *F(fg_ctx *c, int k) { return c->f[k]; }

// ---------------------------------------------------------------- EOS --
// Kernels.cu:167-184
__global__ void k_sound_speed(double *cs, const double *dens, const double *energy, const double *cs_iso, int nr,
                              int ns, int adiabatic, double gamma) {
  CELLS(0, nr)
  if (!adiabatic)
    cs[l] = cs_iso[i];
  else
    cs[l] = sqrt(gamma * (gamma - 1.0) * energy[l] / dens[l]);
}

// Kernels.cu:187-200
__global__ void k_pressure(double *P, const double *dens, const double *cs, const double *energy, int nr, int ns,
                           int adiabatic, double gamma) {
  CELLS(0, nr)
  if (!adiabatic)
    P[l] = dens[l] * cs[l] * cs[l];
  else
    P[l] = (gamma - 1.0) * energy[l];
}

// Kernels.cu:203-213 (MU/R = 1)
__global__ void k_temperature(double *T, const double *dens, const double *P, const double *energy, int nr, int ns,
                              int adiabatic, double gamma) {
  CELLS(0, nr)
  if (!adiabatic)
    T[l] = P[l] / dens[l];
  else
    T[l] = (gamma - 1.0) * energy[l] / dens[l];
}

// Make1DprofileKernel, Kernels.cu:610-623: one ring per block, the ring is
// staged in LDS and summed sequentially by one lane (the reference's j
// order) so the mean is bit-identical.
__global__ void k_ring_mean(double *out, const double *field, int nr, int ns) {
  (void)nr;  // one block per ring: the grid holds the ring count
  extern __shared__ double ring[];
  const int i = blockIdx.x;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) ring[j] = field[(long)i * ns + j];
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int j = 0; j < ns; j++) sum += ring[j];
    out[i] = sum / (double)ns;
  }
}

// FViscosity (Viscosity.cu:33-50) from the ring mean of c_s
__global__ void k_viscosity(double *nu, const double *cs_mean, const double *pow15, const int *cav_on,
                            const double *cav, int nr, int alpha, double alphavisc, double visc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  double v = visc;
  if (alpha) v = alphavisc * cs_mean[i] * cs_mean[i] * pow15[i];
  if (cav_on[i]) v *= cav[i];
  nu[i] = v;
}

// ---------------------------------------------------------- potential --
// FillForcesArraysKernel, Kernels.cu:1226-1253, all planets in one pass
// (the additions keep the reference's order: planet k, then the star
// after planet 0).
__global__ void k_potential(double *pot, const double *Rmed, const double *cosp, const double *sinp,
                            fg_planets_arg P, int nr, int ns, int indirect) {
  CELLS(0, nr)
  const double Rm = Rmed[i];
  const double InvDistance = 1.0 / Rm;
  const double x = Rm * cosp[j];
  const double y = Rm * sinp[j];
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
  pot[l] = acc;
}

// ---------------------------------------------------------- boundaries --
// CorrectVthetaKernel, Kernels.cu:1256-1263
__global__ void k_correct_vtheta(double *vt, const double *Rmed, double domega, int nr, int ns) {
  CELLS(0, nr)
  vt[l] = vt[l] - domega * Rmed[i];
}

// EvanescentBoundaryKernel, Kernels.cu:967-1015 (per-ring constants from
// host tables; lambda, vtheta0 and vrad0 evaluated as in the kernel)
__global__ void k_evanescent(double *vr, double *vt, double *dens, double *energy, fg_dev_tables t, const double *nu,
                             const double *sgacc, double step, double omega_frame, int nr, int ns, int adiabatic,
                             int sg) {
  CELLS(0, nr)
  if (!t.ev_zone[i]) return;
  const double Rm = t.Rmed[i];
  double lambda = t.ev_coef[i] * step / t.ev_period[i];
  double vtheta0 = sg ? sqrt(t.ev_vtheta_arg[i] - Rm * sgacc[i]) : sqrt(t.ev_vtheta_arg[i]);
  vtheta0 -= Rm * omega_frame;
  double vrad0 = -3.0 * nu[i] / Rm * t.vrad0_slope;
  vr[l] = damp_toward(vr[l], lambda, vrad0);
  vt[l] = damp_toward(vt[l], lambda, vtheta0);
  dens[l] = damp_toward(dens[l], lambda, t.SigmaMed[i]);
  if (adiabatic) energy[l] = damp_toward(energy[l], lambda, t.EnergyMed[i]);
}

// OpenBoundaryKernel, Kernels.cu:482-495
__global__ void k_open_boundary(double *vr, double *dens, double *energy, int ns, double sigmamed0) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const int i = 1;
  dens[(i - 1) * ns + j] = dens[i * ns + j];
  energy[(i - 1) * ns + j] = energy[i * ns + j];
  if (vr[(i + 1) * ns + j] > 0.0 || (dens[i * ns + j] < sigmamed0))
    vr[i * ns + j] = 0.0;
  else
    vr[i * ns + j] = vr[(i + 1) * ns + j];
}

// ApplySubKeplerianBoundaryKernel, Kernels.cu:1416-1430 (host passes the
// sqrt arguments; with self-gravity the ring mean of g_r enters)
__global__ void k_sub_keplerian(double *vti, const double *Rmed, const double *sgacc, double arg_in,
                                double arg_out, double omega_frame, int nr, int ns, int sg) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  double vin = sg ? sqrt(arg_in - Rmed[0] * sgacc[0]) : sqrt(arg_in);
  double vout = sg ? sqrt(arg_out - Rmed[nr - 1] * sgacc[nr - 1]) : sqrt(arg_out);
  vti[j] = vin - Rmed[0] * omega_frame;
  vti[(long)(nr - 1) * ns + j] = vout - Rmed[nr - 1] * omega_frame;
}

// --------------------------------------------------------- Substep 1 --
// Substep1Kernel, Kernels.cu:13-53
__global__ void k_substep1(double *vradint, double *vthetaint, const double *P, const double *dens,
                           const double *phi, const double *vr, const double *vt, fg_dev_tables t, double dt,
                           double omega_frame, int nr, int ns) {
  CELLS(0, nr)
  if (i > 0) {
    double gradp = (P[l] - P[l - ns]) * 2.0 / (dens[l] + dens[l - ns]) * t.invdiffRmed[i];
    double gradphi = (phi[l] - phi[l - ns]) * t.invdiffRmed[i];
    double vt2 = vt[l] + vt[JP] + vt[l - ns] + vt[JP - ns];
    vt2 = vt2 / 4.0 + omega_frame * t.Rinf[i];
    vt2 = vt2 * vt2;
    double vr_src = -gradp - gradphi;
    double vradint2 = vr_src + vt2 * t.invRinf[i];
    vradint[l] = vr[l] + dt * vradint2;
  }
  const double invdxtheta = t.invdxtheta[i];
  double gradp = (P[l] - P[JM]) * 2.0 / (dens[l] + dens[JM]) * invdxtheta;
  double gradphi = (phi[l] - phi[JM]) * invdxtheta;
  double v = vt[l] - dt * (gradp + gradphi);
  v += dt * t.supp_torque[i];
  vthetaint[l] = v;
}

// ViscousTermsKernel, Kernels.cu:733-771
__global__ void k_viscous_terms(const double *Vr, const double *Vt, const double *dens, const double *nu,
                                double *DRR, double *DPP, double *DRP, double *divV, double *TRR, double *TPP,
                                double *TRP, fg_dev_tables t, int nr, int ns) {
  CELLS(0, nr)
  const double invdphi = t.invdphi;
  const double onethird = 1.0 / 3.0;
  double drr = (Vr[l + ns] - Vr[l]) * t.invdiffRsup[i];
  double dpp = (Vt[JP] - Vt[l]) * invdphi * t.invRmed[i] + 0.5 * (Vr[l + ns] + Vr[l]) * t.invRmed[i];
  double dv = (Vr[l + ns] * t.Rsup[i] - Vr[l] * t.Rinf[i]) * t.invdiffRsup[i] * t.invRmed[i];
  dv += (Vt[JP] - Vt[l]) * invdphi * t.invRmed[i];
  DRR[l] = drr;
  DPP[l] = dpp;
  divV[l] = dv;
  TRR[l] = 2.0 * dens[l] * nu[i] * (drr - onethird * dv);
  TPP[l] = 2.0 * dens[l] * nu[i] * (dpp - onethird * dv);
  if (i > 0) {
    double drp = 0.5 * (t.Rinf[i] * (Vt[l] * t.invRmed[i] - Vt[l - ns] * t.invRmed[i - 1]) * t.invdiffRmed[i] +
                        (Vr[l] - Vr[JM]) * invdphi * t.invRinf[i]);
    DRP[l] = drp;
    TRP[l] = 2.0 * 0.25 * (dens[l] + dens[l - ns] + dens[JM] + dens[JM - ns]) * nu[i] * drp;
  }
}

// UpdateVelocitiesKernel, Kernels.cu:118-150
__global__ void k_update_velocities(double *vti, double *vri, const double *dens, const double *TRR,
                                    const double *TPP, const double *TRP, fg_dev_tables t, double DeltaT, int nr,
                                    int ns) {
  CELLS(1, nr)
  const double invdphi = t.invdphi;
  if (i < nr - 1)
    vti[l] += DeltaT * t.invRmed[i] *
              ((t.Rsup[i] * TRP[l + ns] - t.Rinf[i] * TRP[l]) * t.invdiffRsup[i] + (TPP[l] - TPP[JM]) * invdphi +
               0.5 * (TRP[l] + TRP[l + ns])) /
              (0.5 * (dens[l] + dens[JM]));
  vri[l] += DeltaT * t.invRinf[i] *
            ((t.Rmed[i] * TRR[l] - t.Rmed[i - 1] * TRR[l - ns]) * t.invdiffRmed[i] + (TRP[JP] - TRP[l]) * invdphi -
             0.5 * (TPP[l] + TPP[l - ns])) /
            (0.5 * (dens[l] + dens[l - ns]));
}

// --------------------------------------------------------- Substep 2 --
// Substep2Kernel, Kernels.cu:409-434
__global__ void k_substep2(double *qr, double *qt, const double *dens, const double *vri, const double *vti, int nr,
                           int ns) {
  CELLS(0, nr)
  qr[l] = art_visc_q(dens[l], vri[l + ns] - vri[l]);
  qt[l] = art_visc_q(dens[l], vti[JP] - vti[l]);
}

// "kernel", Kernels.cu:438-479
__global__ void k_host(double *vrnew, double *vtnew, double *eint, const double *dens, const double *vri,
                       const double *vti, const double *qr, const double *qt, const double *energy, fg_dev_tables t,
                       double dt, int nr, int ns, int adiabatic) {
  CELLS(0, nr)
  if (i > 0) {
    double d = dens[l] + dens[l - ns];
    double densint = qr[l] - qr[l - ns];
    vrnew[l] = vri[l] - dt * 2.0 / d * densint * t.invdiffRmed[i];
  }
  const double invdxtheta = t.invdxtheta[i];
  double dens2 = dens[l] + dens[JM];
  double tempint = (qt[l] - qt[JM]);
  vtnew[l] = vti[l] - dt * 2.0 / dens2 * tempint * invdxtheta;
  if (adiabatic)
    eint[l] = energy[l] - dt * qr[l] * (vri[l + ns] - vri[l]) * t.invdiffRsup[i] -
              dt * qt[l] * (vti[JP] - vti[l]) * invdxtheta;
}

// --------------------------------------------------------- Substep 3 --
// Substep3Kernel, Kernels.cu:56-77
__global__ void k_substep3a(double *qplus, const double *dens, const double *nu, const double *TRR,
                            const double *TRP, const double *TPP, const double *divV, int nr, int ns) {
  CELLS(1, nr)
  if (nu[i] != 0.0) {
    double q = 0.5 / nu[i] / dens[l] * (TRR[l] * TRR[l] + TRP[l] * TRP[l] + TPP[l] * TPP[l]);
    q += (2.0 / 9.0) * nu[i] * dens[l] * divV[l] * divV[l];
    qplus[l] = q;
  } else
    qplus[l] = 0.0;
}

// Substep3Kernel2, Kernels.cu:80-115.  lr01 = log(Rmed0/Rmed1),
// lr12 = log(Rmed1/Rmed2) are host constants; the data-dependent log/exp
// of the power-law extrapolation run on the device.
__global__ void k_substep3b(double *enew, double *qplus, const double *dens, const double *eint, const double *divV,
                            const double *nu, fg_dev_tables t, double dt, double gamma, double lr01, double lr12,
                            int nr, int ns, int cooling) {
  CELLS(0, nr)
  if (i == 0) {
    if (nu[nr - 1] != 0.0)
      qplus[l] = qplus[l + ns] * exp(log(qplus[l + ns] / qplus[l + 2 * ns]) * lr01 / lr12);
    else
      qplus[l] = 0.0;
  }
  double num, den;
  if (!cooling) {
    num = dt * qplus[l] + eint[l];
    den = 1.0 + (gamma - 1.0) * dt * divV[l];
  } else {
    num = t.EnergyMed[i] * dt * dens[l] / t.SigmaMed[i] + t.CoolingTimeMed[i] * eint[l] +
          dt * t.CoolingTimeMed[i] * (qplus[l] - t.QplusMed[i] * dens[l] / t.SigmaMed[i]);
    den = dt + t.CoolingTimeMed[i] + (gamma - 1.0) * dt * t.CoolingTimeMed[i] * divV[l];
  }
  enew[l] = num / den;
}

// --------------------------------------------------------- transport --
// LRMomentaKernel, Kernels.cu:774-787
__global__ void k_lr_momenta(double *rmp, double *rmm, double *tmp_, double *tmm, const double *dens,
                             const double *vr, const double *vt, const double *Rmed, double omega_frame, int nr,
                             int ns) {
  CELLS(0, nr)
  const double Rm = Rmed[i];
  rmp[l] = dens[l] * vr[l + ns];
  rmm[l] = dens[l] * vr[l];
  tmp_[l] = dens[l] * (vt[JP] + Rm * omega_frame) * Rm;
  tmm[l] = dens[l] * (vt[l] + Rm * omega_frame) * Rm;
}

// StarRadKernel, Kernels.cu:800-820
__global__ void k_star_rad_dq(double *dq, const double *q, const double *invdiffRmed, int nr, int ns) {
  CELLS(0, nr)
  if (i == 0 || i == nr - 1)
    dq[l] = 0.0;
  else
    dq[l] = vl_rad_slope(q[l - ns], q[l], q[l + ns], invdiffRmed[i], invdiffRmed[i + 1]);
}

// StarRadKernel2, Kernels.cu:822-838 (Rmed[nrad] is 0; it only ever
// multiplies a zero slope)
__global__ void k_star_rad_q(double *qstar, const double *q, const double *dq, const double *vr, const double *Rmed,
                             double dt, int nr, int ns) {
  CELLS(0, nr)
  double qm = (i > 0) ? q[l - ns] : 0.0;
  double dqm = (i > 0) ? dq[l - ns] : 0.0;
  double rm = (i > 0) ? Rmed[i - 1] : 0.0;
  qstar[l] = vl_rad_star(vr[l], dt, qm, q[l], dqm, dq[l], rm, Rmed[i], Rmed[i + 1], i);
  if (i == 0) {
    qstar[j] = 0.0;
    qstar[(long)ns * nr + j] = 0.0;
  }
}

// DivisePolarGridKernel, Kernels.cu:1018-1025
__global__ void k_divise(double *work, const double *q, const double *densint, long n) {
  long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < n) work[l] = q[l] / (densint[l] + 1e-20);
}

// VanLeerRadialKernel, Kernels.cu:1028-1046
__global__ void k_vl_radial(double *q, double *lost, const double *qrstar, const double *densstar, const double *vr,
                            fg_dev_tables t, double dt, int nr, int ns, int open_inner) {
  CELLS(0, nr)
  const double dtheta = t.dtheta;
  double varq = dt * dtheta * t.Rinf[i] * qrstar[l] * densstar[l] * vr[l];
  varq -= dt * dtheta * t.Rsup[i] * qrstar[l + ns] * densstar[l + ns] * vr[l + ns];
  q[l] += varq * t.invSurf[i];
  if (i == 0 && open_inner) lost[j] = varq;
}

// ComputeResidualsKernel, Kernels.cu:1084-1091
__global__ void k_residual(double *vres, const double *vt, const double *vmed, int nr, int ns) {
  CELLS(0, nr)
  vres[l] = vt[l] - vmed[i];
}

// ComputeConstantResidualKernel, Kernels.cu:1094-1126 (Nshift written by
// lane j == 0 of each ring only)
__global__ void k_const_residual(double *vt, double *vres, int *nshift, int *nosplit, const double *vmed,
                                 const double *invRmed, const double *Rmed, double dt, int nr, int ns, int fast) {
  CELLS(0, nr)
  double maxfrac = fast ? 1.0 : 0.0;
  double invdt = 1.0 / dt;
  double dpinvns = 2.0 * FG_PI / (double)ns;
  double Ntilde = vmed[i] * invRmed[i] * dt * (double)ns / 2.0 / FG_PI;
  double Nround = floor(Ntilde + 0.5);
  long nitemp = (long)Nround;
  if (j == 0) nshift[i] = (int)nitemp;
  double v = (Ntilde - Nround) * Rmed[i] * invdt * dpinvns;
  if (maxfrac < 0.5) {
    if (j == 0) nosplit[i] = 1;
    vres[l] += v;
    v = 0.0;
  } else if (j == 0)
    nosplit[i] = 0;
  vt[l] = v;
}

// StarThetaKernel, Kernels.cu:1129-1148
__global__ void k_star_theta_dq(double *dq, const double *q, const double *invdxtheta, int nr, int ns) {
  CELLS(0, nr)
  dq[l] = vl_theta_slope(q[JM], q[l], q[JP], invdxtheta[i]);
}

// StarThetaKernel2, Kernels.cu:1151-1170
__global__ void k_star_theta_q(double *qstar, const double *q, const double *dq, const double *v,
                               const double *dxtheta, double dt, int nr, int ns) {
  CELLS(0, nr)
  qstar[l] = vl_theta_star(v[l], dt, q[JM], q[l], dq[JM], dq[l], dxtheta[i]);
}

// VanLeerThetaKernel, Kernels.cu:1049-1066
__global__ void k_vl_theta(double *q, const double *qrstar, const double *densstar, const double *v,
                           fg_dev_tables t, double dt, int nr, int ns) {
  CELLS(0, nr)
  double dxrad = (t.Rsup[i] - t.Rinf[i]) * dt;
  double invsurf = t.invSurf[i];
  double varq = dxrad * qrstar[l] * densstar[l] * v[l];
  varq -= dxrad * qrstar[JP] * densstar[JP] * v[JP];
  q[l] += varq * invsurf;
}

// AdvectSHIFTKernel, Kernels.cu:1173-1190
__global__ void k_advect_shift(double *out, const double *in, const int *nshift, int nr, int ns) {
  CELLS(0, nr)
  int ji = (j - nshift[i]) % ns;
  if (ji < 0) ji += ns;
  out[l] = in[(long)i * ns + ji];
}

// ComputeVelocitiesKernel, Kernels.cu:1193-1210
__global__ void k_compute_velocities(double *vr, double *vt, const double *dens, const double *Rmed,
                                     const double *tmp_, const double *tmm, const double *rmp, const double *rmm,
                                     double omega_frame, int nr, int ns) {
  CELLS(0, nr)
  if (i == 0)
    vr[l] = 0.0;
  else
    vr[l] = (rmp[l - ns] + rmm[l]) / (dens[l] + dens[l - ns] + 1e-20);
  vt[l] = (tmp_[JM] + tmm[l]) / (dens[l] + dens[JM] + 1e-15) / Rmed[i] - Rmed[i] * omega_frame;
}

// ------------------------------------------------------------------ CFL --
// ConditionCFLKernel1D + 2D1 + 2D2 (Kernels.cu:1267-1335) for one ring
// per block: sequential ring mean of vtheta (bit-identical), per-cell
// time step, ring minimum (min is order independent).  The same pass sets
// *crash = 1 when a density (or, when energy is not NULL, an energy) of the
// ring is negative or not finite: the check of DetectCrash/CrashKernel
// (SourceEuler.cu:453-465, Kernels.cu:953-964, whose call the reference
// leaves commented out at SourceEuler.cu:266-273), NaN included.
__global__ void k_cfl_ring(double *vmoy, double *newdt, const double *vr, const double *vt, const double *cs,
                           const double *dens, const double *energy, int *crash, fg_dev_tables t, int nr, int ns,
                           int fast) {
  (void)nr;  // one block per ring: the grid holds the ring count
  extern __shared__ double sh[];
  double *ring = sh;               // ns
  double *red = sh + ns;           // blockDim
  __shared__ double s_vmoy;
  const int i = blockIdx.x;
  bool bad = false;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    const long l = (long)i * ns + j;
    ring[j] = vt[l];
    const double d = dens[l];
    bad |= !(d >= 0.0 && d < HUGE_VAL);
    if (energy) {
      const double e = energy[l];
      bad |= !(e >= 0.0 && e < HUGE_VAL);
    }
  }
  if (bad) *crash = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int j = 0; j < ns; j++) m += ring[j];
    m /= (double)ns;
    vmoy[i] = m;
    s_vmoy = m;
  }
  __syncthreads();
  if (i == 0) return;
  const double dxrad = t.Rsup[i] - t.Rinf[i];
  const double dxtheta = t.dxtheta_cfl[i];
  const double mindx = (dxtheta < dxrad) ? dxtheta : dxrad;  // min2
  const double vm = s_vmoy;
  double best = 1e30;
  for (int j = threadIdx.x; j < ns; j += blockDim.x) {
    long l = (long)i * ns + j;
    double vres = fast ? (ring[j] - vm) : ring[j];
    double invdt1 = cs[l] / mindx;
    double invdt2 = fabs(vr[l]) / dxrad;
    double invdt3 = fabs(vres) / dxtheta;
    double dvr = vr[l + ns] - vr[l];
    double dvt = ring[(j + 1) % ns] - ring[j];
    dvr = (dvr >= 0.0) ? 1e-10 : -dvr;
    dvt = (dvt >= 0.0) ? 1e-10 : -dvt;
    double a = dvr / dxrad, b = dvt / dxtheta;
    double invdt4 = (b > a) ? b : a;  // max2
    invdt4 *= 4.0 * 1.41 * 1.41;
    double dtc = 0.5 / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4);
    if (dtc < best) best = dtc;
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && red[threadIdx.x + s] < red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) newdt[i] = red[0];
}

// ConditionCFLKernel2D2 (DT1D) + 2D3, Kernels.cu:1331-1359
__global__ void k_cfl_final(int *cfl, double *dt1d, const double *vmoy, const double *newdt,
                            const double *invRmed, double coef, double DeltaT, int nr) {
  cfl_final(cfl, dt1d, vmoy, newdt, invRmed, coef, DeltaT, nr);
}

// ----------------------------------------------------- entry points --
static fg_status check(fg_ctx *c) {
  if (!c) {
    fg_set_error("NULL context");
    return FG_E_INVALID;
  }
  return FG_OK;
}

extern "C" fg_status fg_sound_speed(fg_ctx *c) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_sound_speed, 0, nr, F(c, FG_SOUNDSPEED), F(c, FG_DENS), F(c, FG_ENERGY), c->t.cs_iso, nr, ns,
         c->p.adiabatic, c->p.adiabatic_index);
  return FG_OK;
}

extern "C" fg_status fg_pressure(fg_ctx *c) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_pressure, 0, nr, F(c, FG_PRESSURE), F(c, FG_DENS), F(c, FG_SOUNDSPEED), F(c, FG_ENERGY), nr, ns,
         c->p.adiabatic, c->p.adiabatic_index);
  return FG_OK;
}

extern "C" fg_status fg_temperature(fg_ctx *c) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_temperature, 0, nr, F(c, FG_TEMPERATURE), F(c, FG_DENS), F(c, FG_PRESSURE), F(c, FG_ENERGY), nr, ns,
         c->p.adiabatic, c->p.adiabatic_index);
  return FG_OK;
}

static fg_status refresh_viscosity(fg_ctx *c, double *nu_out) {
  hipLaunchKernelGGL(k_viscosity, dim3(fg_blocks(c->nr, 256)), dim3(256), 0, c->stream, nu_out, c->cs_mean,
                     c->t.visc_pow15, c->t.visc_cav_on, c->t.visc_cav, c->nr, c->p.viscosity_alpha,
                     c->p.alphaviscosity, c->t.visc_value);
  FG_LAUNCH_CHECK();
  return FG_OK;
}

extern "C" fg_status fg_ring_mean_cs(fg_ctx *c) {
  if (fg_status s = check(c)) return s;
  c->nu_pending = 0;  // nu is current: a fused refresh held back is void
  hipLaunchKernelGGL(k_ring_mean, dim3(c->nr), dim3(256), c->ns * sizeof(double), c->stream, c->cs_mean,
                     F(c, FG_SOUNDSPEED), c->nr, c->ns);
  FG_LAUNCH_CHECK();
  return refresh_viscosity(c, c->nu);
}

// the alpha viscosity of the current sound speed field into nu_out (the
// fused path's refresh at the close of a DT), leaving nu as it is
fg_status fg_viscosity_into(fg_ctx *c, double *nu_out) {
  hipLaunchKernelGGL(k_ring_mean, dim3(c->nr), dim3(256), c->ns * sizeof(double), c->stream, c->cs_mean,
                     F(c, FG_SOUNDSPEED), c->nr, c->ns);
  FG_LAUNCH_CHECK();
  return refresh_viscosity(c, nu_out);
}

extern "C" fg_status fg_cfl(fg_ctx *c, double DeltaT, int *ncfl) {
  if (fg_status s = check(c)) return s;
  if (!ncfl) return FG_E_INVALID;
  const int nr = c->nr, ns = c->ns;
  size_t shm = (size_t)(ns + 256) * sizeof(double);
  hipLaunchKernelGGL(k_cfl_ring, dim3(nr), dim3(256), shm, c->stream, c->Vmoy, c->newDT, F(c, FG_VRAD),
                     F(c, FG_VTHETA), F(c, FG_SOUNDSPEED), F(c, FG_DENS),
                     c->p.adiabatic ? (const double *)F(c, FG_ENERGY) : (const double *)nullptr, c->cfl_out + 1, c->t,
                     nr, ns, c->p.fast_transport);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, c->stream, c->cfl_out, c->DT1D, c->Vmoy, c->newDT,
                     c->t.invRmed, c->t.dt1d_coef, DeltaT, nr);
  FG_LAUNCH_CHECK();
  FG_HIP(hipMemcpyAsync(c->h_cfl, c->cfl_out, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  if (c->h_cfl[1]) {  // reported once; the next pass over a bad state flags it again
    FG_HIP(hipMemsetAsync(c->cfl_out + 1, 0, sizeof(int), c->stream));
    fg_set_error("crash: negative or non-finite density or energy");
    return FG_E_CRASH;
  }
  if (fg_status s = fg_check_substep_results((double)c->h_cfl[0], nullptr)) return s;
  *ncfl = c->h_cfl[0];
  return FG_OK;
}

static fg_status pack_planets(fg_planets_arg *P, const fg_planet_pot *pl, int nb, double ix, double iy) {
  if (nb < 0 || nb > FG_MAX_PLANETS || (nb > 0 && !pl)) {
    fg_set_error("planet count %d outside 0..%d", nb, FG_MAX_PLANETS);
    return FG_E_INVALID;
  }
  P->nb = nb;
  for (int k = 0; k < nb; k++) {
    P->x[k] = pl[k].x;
    P->y[k] = pl[k].y;
    P->m[k] = pl[k].mass;
    P->smooth[k] = pl[k].smooth;
    P->invd3[k] = pl[k].inv_d3;
  }
  P->ix = ix;
  P->iy = iy;
  return FG_OK;
}

extern "C" fg_status fg_potential(fg_ctx *c, const fg_planet_pot *planets, int nb, double ix, double iy) {
  if (fg_status s = check(c)) return s;
  fg_planets_arg P;
  if (fg_status s = pack_planets(&P, planets, nb, ix, iy)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_potential, 0, nr, F(c, FG_POTENTIAL), c->t.Rmed, c->t.cos_pot, c->t.sin_pot, P, nr, ns,
         c->p.indirect_term);
  return FG_OK;
}

extern "C" fg_status fg_correct_vtheta(fg_ctx *c, double domega) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_correct_vtheta, 0, nr, F(c, FG_VTHETA), c->t.Rmed, domega, nr, ns);
  return FG_OK;
}

fg_status fg_nonreflecting_internal(fg_ctx *c);

extern "C" fg_status fg_boundary(fg_ctx *c, double dt, double omega_frame) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  if (c->p.open_inner) {
    double sm0 = c->h_SigmaMed[0];
    hipLaunchKernelGGL(k_open_boundary, dim3(fg_blocks(ns, 256)), dim3(256), 0, c->stream, F(c, FG_VRAD),
                       F(c, FG_DENS), F(c, FG_ENERGY), ns, sm0);
    FG_LAUNCH_CHECK();
  }
  if (c->p.non_reflecting) {
    if (c->p.adiabatic)
      if (fg_status s = fg_sound_speed(c)) return s;
    if (fg_status s = fg_nonreflecting_internal(c)) return s;
  }
  if (c->p.evanescent)
    LAUNCH(k_evanescent, 0, nr, F(c, FG_VRAD), F(c, FG_VTHETA), F(c, FG_DENS), F(c, FG_ENERGY), c->t, c->nu,
           c->sgacc_mean, dt, omega_frame, nr, ns, c->p.adiabatic, c->p.self_gravity);
  return FG_OK;
}

// Viscous stress of (vr, vt) into the TAU fields (Viscosity.cu:52-67).
// With alpha viscosity the ring mean of c_s is refreshed first.
static fg_status viscous_terms(fg_ctx *c, const double *vr, const double *vt) {
  const int nr = c->nr, ns = c->ns;
  if (c->p.viscosity_alpha)
    if (fg_status s = fg_ring_mean_cs(c)) return s;
  LAUNCH(k_viscous_terms, 0, nr, vr, vt, F(c, FG_DENS), c->nu, c->DRR, c->DPP, c->DRP, F(c, FG_DIVV),
         F(c, FG_TAURR), F(c, FG_TAUPP), F(c, FG_TAURP), c->t, nr, ns);
  return FG_OK;
}

fg_status fg_selfgravity_kick(fg_ctx *c, double dt);

extern "C" fg_status fg_substep1(fg_ctx *c, double dt, double omega_frame) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_substep1, 0, nr, F(c, FG_VRADINT), F(c, FG_VTHETAINT), F(c, FG_PRESSURE), F(c, FG_DENS),
         F(c, FG_POTENTIAL), F(c, FG_VRAD), F(c, FG_VTHETA), c->t, dt, omega_frame, nr, ns);
  if (c->p.self_gravity)
    if (fg_status s = fg_selfgravity_kick(c, dt)) return s;
  if (fg_status s = viscous_terms(c, F(c, FG_VRADINT), F(c, FG_VTHETAINT))) return s;
  LAUNCH(k_update_velocities, 1, nr, F(c, FG_VTHETAINT), F(c, FG_VRADINT), F(c, FG_DENS), F(c, FG_TAURR),
         F(c, FG_TAUPP), F(c, FG_TAURP), c->t, dt, nr, ns);
  if (!c->p.evanescent) {
    // SideEuler.cu:194-206: the sqrt arguments are the evanescent
    // vtheta0 arguments of the edge rings (same expression)
    double ain = c->h_vtheta_arg[0], aout = c->h_vtheta_arg[nr - 1];
    hipLaunchKernelGGL(k_sub_keplerian, dim3(fg_blocks(ns, 256)), dim3(256), 0, c->stream, F(c, FG_VTHETAINT),
                       c->t.Rmed, c->sgacc_mean, ain, aout, omega_frame, nr, ns, c->p.self_gravity);
    FG_LAUNCH_CHECK();
  }
  return FG_OK;
}

extern "C" fg_status fg_substep2(fg_ctx *c, double dt) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_substep2, 0, nr, F(c, FG_DENSINT), F(c, FG_TEMPERINT), F(c, FG_DENS), F(c, FG_VRADINT),
         F(c, FG_VTHETAINT), nr, ns);
  LAUNCH(k_host, 0, nr, F(c, FG_VRADNEW), F(c, FG_VTHETANEW), F(c, FG_ENERGYINT), F(c, FG_DENS), F(c, FG_VRADINT),
         F(c, FG_VTHETAINT), F(c, FG_DENSINT), F(c, FG_TEMPERINT), F(c, FG_ENERGY), c->t, dt, nr, ns,
         c->p.adiabatic);
  const size_t bytes = c->S * sizeof(double);
  FG_HIP(hipMemcpyAsync(F(c, FG_VRAD), F(c, FG_VRADNEW), bytes, hipMemcpyDeviceToDevice, c->stream));
  FG_HIP(hipMemcpyAsync(F(c, FG_VTHETA), F(c, FG_VTHETANEW), bytes, hipMemcpyDeviceToDevice, c->stream));
  return FG_OK;
}


extern "C" fg_status fg_substep3(fg_ctx *c, double dt) {
  if (fg_status s = check(c)) return s;
  if (!c->p.adiabatic) return FG_OK;
  const int nr = c->nr, ns = c->ns;
  if (fg_status s = viscous_terms(c, F(c, FG_VRAD), F(c, FG_VTHETA))) return s;
  LAUNCH(k_substep3a, 1, nr, F(c, FG_QPLUS), F(c, FG_DENS), c->nu, F(c, FG_TAURR), F(c, FG_TAURP), F(c, FG_TAUPP),
         F(c, FG_DIVV), nr, ns);
  const double lr01 = c->lr01, lr12 = c->lr12;
  LAUNCH(k_substep3b, 0, nr, F(c, FG_ENERGYNEW), F(c, FG_QPLUS), F(c, FG_DENS), F(c, FG_ENERGYINT), F(c, FG_DIVV),
         c->nu, c->t, dt, c->p.adiabatic_index, lr01, lr12, nr, ns, c->p.cooling);
  FG_HIP(hipMemcpyAsync(F(c, FG_ENERGY), F(c, FG_ENERGYNEW), c->S * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream));
  return FG_OK;
}

fg_status fg_reduce_array_internal(fg_ctx *c, const double *in, int n, double *result_dev);

// ComputeStarRad (TransportEuler.cu:81-95) of q into qstar
static fg_status star_rad(fg_ctx *c, const double *q, double *qstar, double dt) {
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_star_rad_dq, 0, nr, c->dq, q, c->t.invdiffRmed, nr, ns);
  LAUNCH(k_star_rad_q, 0, nr, qstar, q, c->dq, F(c, FG_VRAD), c->t.Rmed, dt, nr, ns);
  return FG_OK;
}

// VanLeerRadial (TransportEuler.cu:106-156)
static fg_status van_leer_radial(fg_ctx *c, double *q, double dt) {
  const int nr = c->nr, ns = c->ns;
  long n = (long)nr * ns;
  hipLaunchKernelGGL(k_divise, dim3(fg_blocks(n, TPB)), dim3(TPB), 0, c->stream, c->Work, q, F(c, FG_DENSINT), n);
  FG_LAUNCH_CHECK();
  if (fg_status s = star_rad(c, c->Work, c->QRStar, dt)) return s;
  LAUNCH(k_vl_radial, 0, nr, q, c->LostByDisk, c->QRStar, c->DensStar, F(c, FG_VRAD), c->t, dt, nr, ns,
         c->p.open_inner);
  return FG_OK;
}

// ComputeStarTheta (TransportEuler.cu:261-269)
static fg_status star_theta(fg_ctx *c, const double *q, const double *v, double *qstar, double dt) {
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_star_theta_dq, 0, nr, c->dq, q, c->t.invdxtheta, nr, ns);
  LAUNCH(k_star_theta_q, 0, nr, qstar, q, c->dq, v, c->t.dxtheta, dt, nr, ns);
  return FG_OK;
}

// VanLeerTheta (TransportEuler.cu:249-257)
static fg_status van_leer_theta(fg_ctx *c, const double *v, double *q, double dt) {
  const int nr = c->nr, ns = c->ns;
  long n = (long)nr * ns;
  hipLaunchKernelGGL(k_divise, dim3(fg_blocks(n, TPB)), dim3(TPB), 0, c->stream, c->Work, q, F(c, FG_DENSINT), n);
  FG_LAUNCH_CHECK();
  if (fg_status s = star_theta(c, c->Work, v, c->QRStar, dt)) return s;
  LAUNCH(k_vl_theta, 0, nr, q, c->QRStar, c->DensStar, v, c->t, dt, nr, ns);
  return FG_OK;
}

// QuantitiesAdvection (TransportEuler.cu:229-245)
static fg_status quantities_advection(fg_ctx *c, const double *v, double dt) {
  if (fg_status s = star_theta(c, F(c, FG_DENS), v, c->DensStar, dt)) return s;
  FG_HIP(hipMemcpyAsync(F(c, FG_DENSINT), F(c, FG_DENS), c->S * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  const int order[4] = {FG_RADMOMP, FG_RADMOMM, FG_THETAMOMP, FG_THETAMOMM};
  for (int k = 0; k < 4; k++)
    if (fg_status s = van_leer_theta(c, v, F(c, order[k]), dt)) return s;
  if (c->p.adiabatic)
    if (fg_status s = van_leer_theta(c, v, F(c, FG_ENERGY), dt)) return s;
  return van_leer_theta(c, v, F(c, FG_DENS), dt);
}

static fg_status advect_shift(fg_ctx *c, double *q) {
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_advect_shift, 0, nr, c->TempShift, q, c->Nshift, nr, ns);
  FG_HIP(hipMemcpyAsync(q, c->TempShift, c->S * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return FG_OK;
}

extern "C" fg_status fg_transport(fg_ctx *c, double dt, double omega_frame, double *lost_mass) {
  if (fg_status s = check(c)) return s;
  const int nr = c->nr, ns = c->ns;
  LAUNCH(k_lr_momenta, 0, nr, F(c, FG_RADMOMP), F(c, FG_RADMOMM), F(c, FG_THETAMOMP), F(c, FG_THETAMOMM),
         F(c, FG_DENS), F(c, FG_VRAD), F(c, FG_VTHETA), c->t.Rmed, omega_frame, nr, ns);
  // OneWindRad
  if (fg_status s = star_rad(c, F(c, FG_DENS), c->DensStar, dt)) return s;
  FG_HIP(hipMemcpyAsync(F(c, FG_DENSINT), F(c, FG_DENS), c->S * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  const int order[4] = {FG_RADMOMP, FG_RADMOMM, FG_THETAMOMP, FG_THETAMOMM};
  for (int k = 0; k < 4; k++)
    if (fg_status s = van_leer_radial(c, F(c, order[k]), dt)) return s;
  if (c->p.adiabatic)
    if (fg_status s = van_leer_radial(c, F(c, FG_ENERGY), dt)) return s;
  if (fg_status s = van_leer_radial(c, F(c, FG_DENS), dt)) return s;
  if (fg_status s = fg_reduce_array_internal(c, c->LostByDisk, ns, c->red_out + 2)) return s;
  // OneWindTheta
  hipLaunchKernelGGL(k_ring_mean, dim3(nr), dim3(256), ns * sizeof(double), c->stream, c->VMed, F(c, FG_VTHETA), nr,
                     ns);
  FG_LAUNCH_CHECK();
  LAUNCH(k_residual, 0, nr, F(c, FG_VTHETARES), F(c, FG_VTHETA), c->VMed, nr, ns);
  LAUNCH(k_const_residual, 0, nr, F(c, FG_VTHETA), F(c, FG_VTHETARES), c->Nshift, c->NoSplit, c->VMed,
         c->t.invRmed, c->t.Rmed, dt, nr, ns, c->p.fast_transport);
  if (fg_status s = quantities_advection(c, F(c, FG_VTHETARES), dt)) return s;
  if (fg_status s = quantities_advection(c, F(c, FG_VTHETA), dt)) return s;
  for (int k = 0; k < 4; k++)
    if (fg_status s = advect_shift(c, F(c, order[k]))) return s;
  if (c->p.adiabatic)
    if (fg_status s = advect_shift(c, F(c, FG_ENERGY))) return s;
  if (fg_status s = advect_shift(c, F(c, FG_DENS))) return s;
  LAUNCH(k_compute_velocities, 0, nr, F(c, FG_VRAD), F(c, FG_VTHETA), F(c, FG_DENS), c->t.Rmed, F(c, FG_THETAMOMP),
         F(c, FG_THETAMOMM), F(c, FG_RADMOMP), F(c, FG_RADMOMM), omega_frame, nr, ns);
  if (lost_mass) {
    FG_HIP(hipMemcpyAsync(c->h_out + 2, c->red_out + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    *lost_mass = c->h_out[2];
  }
  return FG_OK;
}
