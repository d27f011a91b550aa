// Context lifetime, uploads/downloads and field access of libfargo_hip.so.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "fg_internal.h"

static thread_local char g_fg_err[1024] = "";

void fg_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_fg_err, sizeof g_fg_err, fmt, ap);
  va_end(ap);
}

extern "C" const char *fg_last_error(void) { return g_fg_err; }

template <typename T>
static fg_status dev_alloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    fg_set_error("hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    return FG_E_NOMEM;
  }
  e = hipMemset(*p, 0, n * sizeof(T));
  if (e != hipSuccess) {
    fg_set_error("hipMemset: %s", hipGetErrorString(e));
    return FG_E_HIP;
  }
  return FG_OK;
}

template <typename T>
static fg_status dev_upload(T **p, const T *host, size_t n) {
  fg_status st = dev_alloc(p, n);
  if (st) return st;
  if (host) FG_HIP(hipMemcpy(*p, host, n * sizeof(T), hipMemcpyHostToDevice));
  return FG_OK;
}

#define TRY(x)                  \
  do {                          \
    fg_status st_ = (x);        \
    if (st_) {                  \
      fg_destroy(c);            \
      return st_;               \
    }                           \
  } while (0)

extern "C" fg_status fg_create(const fg_params *p, const fg_tables *t, int device, fg_ctx **out) {
  if (!p || !t || !out) {
    fg_set_error("fg_create: NULL argument");
    return FG_E_INVALID;
  }
  *out = nullptr;
  if (p->nrad < 3 || p->nsec < 2) {
    fg_set_error("fg_create: grid %d x %d too small (NRAD >= 3, NSEC >= 2)", p->nrad, p->nsec);
    return FG_E_INVALID;
  }
  if (p->nsec > 8192) {
    fg_set_error("fg_create: NSEC = %d exceeds the 8192 sectors a ring kernel stages in LDS", p->nsec);
    return FG_E_UNSUPPORTED;
  }
  int ndev = 0;
  FG_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    fg_set_error("fg_create: device %d not present (%d visible)", device, ndev);
    return FG_E_INVALID;
  }
  FG_HIP(hipSetDevice(device));
  // this device's copy of the allocation-free entry point's scratch array;
  // a failure is reported by that entry point (fg_disk_force_hill_noalloc),
  // the context itself does not need it
  (void)fg_noalloc_setup(device);
  fg_ctx *c = new fg_ctx();
  c->tile_src_th = 16;
  c->tile_trad_band = 16;
  c->fp32_pair_hill = 1.0;
  c->p = *p;
  c->nr = p->nrad;
  c->ns = p->nsec;
  c->S = (long)(c->nr + 1) * c->ns;
  c->device = device;
  hipError_t e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    fg_set_error("hipStreamCreate: %s", hipGetErrorString(e));
    delete c;
    return FG_E_HIP;
  }
  const size_t R = (size_t)c->nr + 1, NS = (size_t)c->ns, S = (size_t)c->S;
  fg_dev_tables &d = c->t;
  TRY(dev_upload(&d.Radii, t->Radii, R));
  TRY(dev_upload(&d.Rinf, t->Rinf, R));
  TRY(dev_upload(&d.Rsup, t->Rsup, R));
  TRY(dev_upload(&d.Rmed, t->Rmed, R));
  TRY(dev_upload(&d.Surf, t->Surf, R));
  TRY(dev_upload(&d.invRinf, t->invRinf, R));
  TRY(dev_upload(&d.invSurf, t->invSurf, R));
  TRY(dev_upload(&d.invRmed, t->invRmed, R));
  TRY(dev_upload(&d.invdiffRsup, t->invdiffRsup, R));
  TRY(dev_upload(&d.invdiffRmed, t->invdiffRmed, R));
  TRY(dev_upload(&d.SigmaMed, t->SigmaMed, R));
  TRY(dev_upload(&d.SigmaInf, t->SigmaInf, R));
  TRY(dev_upload(&d.EnergyMed, t->EnergyMed, R));
  TRY(dev_upload(&d.CoolingTimeMed, t->CoolingTimeMed, R));
  TRY(dev_upload(&d.QplusMed, t->QplusMed, R));
  TRY(dev_upload(&d.cs_iso, t->cs_iso, R));
  TRY(dev_upload(&d.supp_torque, t->supp_torque, R));
  TRY(dev_upload(&d.dxtheta, t->dxtheta, R));
  TRY(dev_upload(&d.invdxtheta, t->invdxtheta, R));
  TRY(dev_upload(&d.dxtheta_cfl, t->dxtheta_cfl, R));
  TRY(dev_upload(&d.ev_zone, t->ev_zone, R));
  TRY(dev_upload(&d.ev_coef, t->ev_coef, R));
  TRY(dev_upload(&d.ev_period, t->ev_period, R));
  TRY(dev_upload(&d.ev_vtheta_arg, t->ev_vtheta_arg, R));
  TRY(dev_upload(&d.visc_const, t->visc_const, R));
  TRY(dev_upload(&d.visc_pow15, t->visc_pow15, R));
  TRY(dev_upload(&d.visc_cav_on, t->visc_cav_on, R));
  TRY(dev_upload(&d.visc_cav, t->visc_cav, R));
  TRY(dev_upload(&d.cos_cell, t->cos_cell, NS));
  TRY(dev_upload(&d.sin_cell, t->sin_cell, NS));
  TRY(dev_upload(&d.cos_pot, t->cos_pot, NS));
  TRY(dev_upload(&d.sin_pot, t->sin_pot, NS));
  c->h_Rmed.assign(t->Rmed, t->Rmed + R);
  c->h_Rinf.assign(t->Rinf, t->Rinf + R);
  c->h_SigmaMed.assign(t->SigmaMed, t->SigmaMed + R);
  c->h_EnergyMed.assign(t->EnergyMed, t->EnergyMed + R);
  c->h_vtheta_arg.assign(t->ev_vtheta_arg, t->ev_vtheta_arg + R);
  c->lr01 = log(t->Rmed[0] / t->Rmed[1]);
  c->lr12 = log(t->Rmed[1] / t->Rmed[2]);
  d.dtheta = t->dtheta;
  d.invdphi = t->invdphi;
  d.dt1d_coef = t->dt1d_coef;
  d.vrad0_slope = t->vrad0_slope;
  d.visc_value = t->visc_value;
  if (p->self_gravity) {
    size_t K = 2 * (size_t)c->nr * NS;
    TRY(dev_upload(&d.sg_Kr, t->sg_Kr, K));
    TRY(dev_upload(&d.sg_Kt, t->sg_Kt, K));
    // per-ring constants of ComputeSgAccKernel / ComputeFFTKernel
    // (Kernels.cu:859-861, 902-906): only sqrt and /, evaluated here.
    std::vector<double> nr_(R, 0.0), nt_(R, 0.0), sq(R, 0.0), ra(R, 0.0);
    for (int i = 0; i < c->nr; i++) {
      double divRmed = t->Rmed[i] / t->Rmed[0];
      double normaccr = p->sg_rstep * p->sg_tstep / ((double)(2 * c->nr) * (double)c->ns);
      double normacct = normaccr;
      normaccr /= sqrt(divRmed);
      normacct /= (divRmed * sqrt(divRmed));
      nr_[i] = normaccr;
      nt_[i] = normacct;
      sq[i] = sqrt(t->Rmed[i] / t->Rmed[0]);
      ra[i] = divRmed;
    }
    TRY(dev_upload(&d.sg_norm_r, nr_.data(), R));
    TRY(dev_upload(&d.sg_norm_t, nt_.data(), R));
    TRY(dev_upload(&d.sg_sqrt_ratio, sq.data(), R));
    TRY(dev_upload(&d.sg_ratio, ra.data(), R));
    size_t C = 2 * K;  // interleaved complex
    TRY(dev_alloc(&c->sg_Sr, C));
    TRY(dev_alloc(&c->sg_St, C));
    TRY(dev_alloc(&c->sg_Krc, C));
    TRY(dev_alloc(&c->sg_Ktc, C));
    TRY(dev_alloc(&c->sg_Gr, C));
    TRY(dev_alloc(&c->sg_Gt, C));
    TRY(dev_alloc(&c->sg_tmp, C));
  }
  for (int k = 0; k < FG_NFIELDS; k++) TRY(dev_alloc(&c->f[k], S));
  TRY(dev_alloc(&c->DRR, S));
  TRY(dev_alloc(&c->DPP, S));
  TRY(dev_alloc(&c->DRP, S));
  TRY(dev_alloc(&c->Work, S));
  TRY(dev_alloc(&c->QRStar, S));
  TRY(dev_alloc(&c->DensStar, S));
  TRY(dev_alloc(&c->dq, S));
  TRY(dev_alloc(&c->TempShift, S));
  TRY(dev_alloc(&c->Vradial, S));
  TRY(dev_alloc(&c->Vazimutal, S));
  TRY(dev_alloc(&c->Vmoy, R));
  TRY(dev_alloc(&c->newDT, R));
  TRY(dev_alloc(&c->DT1D, R));
  TRY(dev_alloc(&c->VMed, R));
  TRY(dev_alloc(&c->nu, R));
  TRY(dev_alloc(&c->cs_mean, R));
  TRY(dev_alloc(&c->sgacc_mean, R));
  TRY(dev_alloc(&c->sgacc_prev, R));
  TRY(dev_alloc(&c->nu_next, R));
  TRY(dev_alloc(&c->Nshift, R));
  TRY(dev_alloc(&c->NoSplit, R));
  TRY(dev_alloc(&c->LostByDisk, NS));
  TRY(dev_alloc(&c->red_partial, 2 * FG_RED_BLOCKS));
  TRY(dev_alloc(&c->red_out, 8));
  TRY(dev_alloc(&c->cfl_out, 4));
  if (hipHostMalloc((void **)&c->h_out, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->h_cfl, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    fg_set_error("hipHostMalloc of the result mirror failed");
    fg_destroy(c);
    return FG_E_NOMEM;
  }
  // initial viscosity profile (constant viscosity; alpha viscosity is
  // refreshed by fg_ring_mean_cs)
  {
    std::vector<double> nu(R, 0.0);
    for (int i = 0; i < c->nr; i++) {
      double v = t->visc_value;
      if (t->visc_cav_on[i]) v *= t->visc_cav[i];
      nu[i] = v;
    }
    FG_HIP(hipMemcpy(c->nu, nu.data(), R * sizeof(double), hipMemcpyHostToDevice));
  }
  // isothermal sound speed field: constant in time (Kernels.cu:179 is
  // only evaluated at initialisation for a locally isothermal disk)
  {
    std::vector<double> cs(S, 0.0);
    for (int i = 0; i < c->nr; i++)
      for (int j = 0; j < c->ns; j++) cs[(size_t)i * c->ns + j] = t->cs_iso[i];
    FG_HIP(hipMemcpy(c->f[FG_SOUNDSPEED], cs.data(), S * sizeof(double), hipMemcpyHostToDevice));
  }
  *out = c;
  return FG_OK;
}

extern "C" fg_status fg_destroy(fg_ctx *c) {
  if (!c) return FG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  fg_sg_release(c);
  void *ptrs[] = {c->t.Radii,      c->t.Rinf,       c->t.Rsup,        c->t.Rmed,         c->t.Surf,
                  c->t.invRinf,    c->t.invSurf,    c->t.invRmed,     c->t.invdiffRsup,  c->t.invdiffRmed,
                  c->t.SigmaMed,   c->t.SigmaInf,   c->t.EnergyMed,   c->t.CoolingTimeMed, c->t.QplusMed,
                  c->t.cs_iso,     c->t.supp_torque, c->t.dxtheta,    c->t.invdxtheta,   c->t.dxtheta_cfl,
                  c->t.ev_zone,    c->t.ev_coef,    c->t.ev_period,   c->t.ev_vtheta_arg, c->t.visc_const,
                  c->t.visc_pow15, c->t.visc_cav,   c->t.visc_cav_on, c->t.cos_cell,     c->t.sin_cell,
                  c->t.cos_pot,    c->t.sin_pot,    c->t.sg_Kr,       c->t.sg_Kt,        c->t.sg_norm_r,
                  c->t.sg_norm_t,  c->t.sg_sqrt_ratio, c->t.sg_ratio, c->DRR,            c->DPP,
                  c->DRP,          c->Work,         c->QRStar,        c->DensStar,       c->dq,
                  c->TempShift,    c->Vradial,      c->Vazimutal,     c->Vmoy,           c->newDT,
                  c->DT1D,         c->VMed,         c->nu,            c->cs_mean,        c->sgacc_mean,
                  c->sgacc_prev,   c->nu_next,
                  c->Nshift,       c->NoSplit,      c->LostByDisk,    c->red_partial,    c->red_out,
                  c->cfl_out,      c->sg_Sr,        c->sg_St,         c->sg_Krc,         c->sg_Ktc,
                  c->sg_Gr,        c->sg_Gt,        c->sg_tmp,        c->fu_ring,        c->fu_fpart,
                  c->hill_partial, c->hill_out};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (int k = 0; k < FG_NFIELDS; k++)
    if (c->f[k]) (void)hipFree(c->f[k]);
  for (int k = 0; k < 8; k++)
    if (c->f32[k]) (void)hipFree(c->f32[k]);
  for (int k = 0; k < 4; k++)
    if (c->fu_mom[k]) (void)hipFree(c->fu_mom[k]);
  if (c->h_out) (void)hipHostFree(c->h_out);
  if (c->h_fused) (void)hipHostFree(c->h_fused);
  if (c->h_cfl) (void)hipHostFree(c->h_cfl);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FG_OK;
}

extern "C" fg_status fg_sync(fg_ctx *c) {
  if (!c) return FG_E_INVALID;
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}

static fg_status put(fg_ctx *c, double *dst, const double *src) {
  if (src) FG_HIP(hipMemcpyAsync(dst, src, c->S * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return FG_OK;
}
static fg_status get(fg_ctx *c, double *dst, const double *src) {
  if (dst) FG_HIP(hipMemcpyAsync(dst, src, c->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return FG_OK;
}

extern "C" fg_status fg_upload_state(fg_ctx *c, const double *dens, const double *vrad, const double *vtheta,
                                     const double *energy) {
  if (!c) return FG_E_INVALID;
  fg_status st;
  if ((st = put(c, c->f[FG_DENS], dens))) return st;
  if ((st = put(c, c->f[FG_VRAD], vrad))) return st;
  if ((st = put(c, c->f[FG_VTHETA], vtheta))) return st;
  if ((st = put(c, c->f[FG_ENERGY], energy))) return st;
  // a new state: crash flags raised by passes over the old one no longer apply
  FG_HIP(hipMemsetAsync(c->cfl_out + 1, 0, sizeof(int), c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  if (c->h_fused) *(volatile int *)(c->h_fused + FG_FUSED_CRASH_SLOT) = 0;
  return FG_OK;
}

extern "C" fg_status fg_download_state(fg_ctx *c, double *dens, double *vrad, double *vtheta, double *energy,
                                       double *temperature, double *qplus) {
  if (!c) return FG_E_INVALID;
  fg_status st;
  if ((st = get(c, dens, c->f[FG_DENS]))) return st;
  if ((st = get(c, vrad, c->f[FG_VRAD]))) return st;
  if ((st = get(c, vtheta, c->f[FG_VTHETA]))) return st;
  if ((st = get(c, energy, c->f[FG_ENERGY]))) return st;
  if ((st = get(c, temperature, c->f[FG_TEMPERATURE]))) return st;
  if ((st = get(c, qplus, c->f[FG_QPLUS]))) return st;
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}

extern "C" fg_status fg_get_field(fg_ctx *c, int field, double *host) {
  if (!c || !host || field < 0 || field >= FG_NFIELDS) {
    fg_set_error("fg_get_field: bad argument");
    return FG_E_INVALID;
  }
  FG_HIP(hipMemcpyAsync(host, c->f[field], c->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}

extern "C" fg_status fg_set_field(fg_ctx *c, int field, const double *host) {
  if (!c || !host || field < 0 || field >= FG_NFIELDS) {
    fg_set_error("fg_set_field: bad argument");
    return FG_E_INVALID;
  }
  FG_HIP(hipMemcpyAsync(c->f[field], host, c->S * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}

extern "C" fg_status fg_get_rings(fg_ctx *c, double *vmoy, double *vmed, double *viscosity, double *cs_mean,
                                  double *sgacc_mean) {
  if (!c) return FG_E_INVALID;
  size_t n = (size_t)c->nr * sizeof(double);
  if (vmoy) FG_HIP(hipMemcpyAsync(vmoy, c->Vmoy, n, hipMemcpyDeviceToHost, c->stream));
  if (vmed) FG_HIP(hipMemcpyAsync(vmed, c->VMed, n, hipMemcpyDeviceToHost, c->stream));
  if (viscosity) FG_HIP(hipMemcpyAsync(viscosity, c->nu, n, hipMemcpyDeviceToHost, c->stream));
  if (cs_mean) FG_HIP(hipMemcpyAsync(cs_mean, c->cs_mean, n, hipMemcpyDeviceToHost, c->stream));
  if (sgacc_mean) FG_HIP(hipMemcpyAsync(sgacc_mean, c->sgacc_mean, n, hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}

extern "C" fg_status fg_get_nshift(fg_ctx *c, int *nshift) {
  if (!c || !nshift) return FG_E_INVALID;
  FG_HIP(hipMemcpyAsync(nshift, c->Nshift, (size_t)c->nr * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}
