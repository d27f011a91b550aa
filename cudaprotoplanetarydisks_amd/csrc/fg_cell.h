// Per-cell arithmetic of the FARGO substep, shared by the per-routine
// kernels (fg_steps.hip) and the fused ring-band kernels (fg_fused.hip).
//
// Each function evaluates one reference expression in the reference's own
// left-to-right order (the build uses -ffp-contract=off), so both paths
// round exactly like the FP64 oracle.  Indices: l = i*ns + j, row-major.
#pragma once

#include <hip/hip_runtime.h>

#define FG_DEV __device__ __forceinline__

// n / d for many n with one d: the compiler's FP64 division (v_div_scale,
// v_rcp_f64 and two Newton steps on the scaled denominator, the corrected
// quotient, v_div_fmas, v_div_fixup) with the denominator's half done once.
// Neither operand is scaled and the last fma stands for v_div_fmas, so the
// result is bit-identical to n / d whenever v_div_scale would scale
// neither: d and 1/d normal, and n zero, infinite, NaN or finite with
// 2^-969 <= |n| and |n / d| < 2^767 (v_div_fixup returns the same signed
// zero, infinity or NaN as the full sequence).
struct recip64 {
  double d, r;
  FG_DEV explicit recip64(double den) : d(den) {
    r = __builtin_amdgcn_rcp(den);
    double t = fma(-den, r, 1.0);
    r = fma(r, t, r);
    t = fma(-den, r, 1.0);
    r = fma(r, t, r);
  }
  FG_DEV recip64(double den, double refined) : d(den), r(refined) {}  // r of an earlier recip64(den)
  FG_DEV double div(double n) const {
    const double q0 = n * r;
    const double e = fma(-d, q0, n);
    return __builtin_amdgcn_div_fixup(fma(e, r, q0), d, n);
  }
};

// van Leer radial slope, Kernels.cu:808-818
FG_DEV double vl_rad_slope(double qm, double q0, double qp, double idrm_i, double idrm_ip) {
  double dqm = (q0 - qm) * idrm_i;
  double dqp = (qp - q0) * idrm_ip;
  if (dqp * dqm > 0.0) return 2.0 * dqp * dqm / (dqp + dqm);
  return 0.0;
}

// upwind interface value at Rinf[i], Kernels.cu:828-833.  Callers handle
// the i == 0 and i == nrad rows, which the reference zeroes.
FG_DEV double vl_rad_star(double vr, double dt, double q_im1, double q_i, double dq_im1, double dq_i,
                          double rmed_im1, double rmed_i, double rmed_ip1, int i) {
  if (vr > 0.0 && i > 0) return q_im1 + (rmed_i - rmed_im1 - vr * dt) * 0.5 * dq_im1;
  return q_i - (rmed_ip1 - rmed_i + vr * dt) * 0.5 * dq_i;
}

// van Leer azimuthal slope, Kernels.cu:1140-1146
FG_DEV double vl_theta_slope(double qm, double q0, double qp, double invdxtheta) {
  double dqm = (q0 - qm);
  double dqp = (qp - q0);
  if (dqp * dqm > 0.0) return dqp * dqm / (dqp + dqm) * invdxtheta;
  return 0.0;
}

// upwind interface value at the j-1/2 face, Kernels.cu:1164-1168
FG_DEV double vl_theta_star(double v, double dt, double q_jm1, double q_j, double dq_jm1, double dq_j,
                            double dxtheta) {
  double ksi = v * dt;
  if (ksi > 0.0) return q_jm1 + (dxtheta - ksi) * dq_jm1;
  return q_j - (dxtheta + ksi) * dq_j;
}

// von Neumann-Richtmyer artificial pressure, Kernels.cu:420-432
FG_DEV double art_visc_q(double dens, double dv) {
  if (dv < 0.0) return 1.41 * 1.41 * dens * dv * dv;
  return 0.0;
}

// evanescent relaxation, Kernels.cu:1006-1012
FG_DEV double damp_toward(double q, double lambda, double q0) { return (q + lambda * q0) / (1.0 + lambda); }

// ---- FP32 arithmetic of the bulk-precision mode ----------------------
// Same formulas as above in single precision.  Quotients use the hardware
// reciprocal (1 ulp): the FP32 mode is held to an accuracy bound against
// the FP64 solution, not to bit equality.
FG_DEV float rcp_f(float b) { return __builtin_amdgcn_rcpf(b); }

FG_DEV float vl_rad_slope_f(float qm, float q0, float qp, float idrm_i, float idrm_ip) {
  float dqm = (q0 - qm) * idrm_i;
  float dqp = (qp - q0) * idrm_ip;
  if (dqp * dqm > 0.0f) return 2.0f * dqp * dqm * rcp_f(dqp + dqm);
  return 0.0f;
}

// dm = Rmed[i] - Rmed[i-1], dp = Rmed[i+1] - Rmed[i], formed in FP64 by
// the caller (the ring spacing is ~1e-3 of the radius)
FG_DEV float vl_rad_star_f(float vr, float dt, float q_im1, float q_i, float dq_im1, float dq_i, float dm,
                           float dp) {
  if (vr > 0.0f) return q_im1 + (dm - vr * dt) * 0.5f * dq_im1;
  return q_i - (dp + vr * dt) * 0.5f * dq_i;
}

FG_DEV float vl_theta_slope_f(float qm, float q0, float qp, float invdxtheta) {
  float dqm = (q0 - qm);
  float dqp = (qp - q0);
  if (dqp * dqm > 0.0f) return dqp * dqm * rcp_f(dqp + dqm) * invdxtheta;
  return 0.0f;
}

FG_DEV float vl_theta_star_f(float v, float dt, float q_jm1, float q_j, float dq_jm1, float dq_j, float dxtheta) {
  float ksi = v * dt;
  if (ksi > 0.0f) return q_jm1 + (dxtheta - ksi) * dq_jm1;
  return q_j - (dxtheta + ksi) * dq_j;
}

FG_DEV float art_visc_q_f(float dens, float dv) {
  if (dv < 0.0f) return (1.41f * 1.41f) * dens * dv * dv;
  return 0.0f;
}
