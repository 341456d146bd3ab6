// statics.hip — imaginary-time relaxation of He_statics: STEPRKR / STEPPCR, the
// log-psi steppers STEPRKRL / STEPPCRL, steepest descent (evolo + ironing), the
// impurity wave function (evolox), and the driver entry points around them.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "internal.h"

using namespace hedft;

namespace hedft {

// ---- statics: hpsi = -(lap)*h2o2m4 + pot4*psi and sum(psi*hpsi) (steprkr.f90:52-53)
struct OutHpsiR {
  static constexpr int NR = 1;
  double* partials;
  const double *lapxy, *psi, *pot4;
  double* hpsi;
  double h2o2m4;
  __device__ __forceinline__ void store(size_t i, int, int, int, double dz, double* red) const {
    const double p = psi[i];
    const double hp = -(lapxy[i] + dz) * h2o2m4 + pot4[i] * p;
    hpsi[i] = hp;
    red[0] += p * hp;
  }
};

// ---- statics elementwise kernels (grid-stride, fixed grid for deterministic sums)

// one RK-Gill stage of STEPRKR after the mean (steprkr.f90:54-73), sum(psi^2)
__global__ void __launch_bounds__(256) k_rkr_stage(long long n, double* psi, const double* hpsi, double* q,
                                                   double* timec, double* po0, double* po1, double* po2,
                                                   double* ho0, double* ho1, const double* scal, double a,
                                                   double b, double cc, double deltat, double dtmin,
                                                   double sdmu, double dmumax, double psimin, int first,
                                                   double* partials) {
  double red[1] = {0.0};
  const double mu = scal[S_MU];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double p = psi[i], hp = hpsi[i];
    if (first) timec[i] = deltat / (1.0 + exp((fabs(hp / p) - dmumax) / sdmu)) + dtmin;
    const double s4 = -(hp - mu * p);
    const double s1 = a * (s4 - b * q[i]);
    q[i] = q[i] + 3.0 * s1 - cc * s4;
    if (first) {
      ho1[i] = ho0[i];
      ho0[i] = s4;
      po2[i] = po1[i];
      po1[i] = po0[i];
      po0[i] = p;
    }
    double np = p + timec[i] * s1;
    if (psimin > 0.0 && np < psimin) np = psimin;
    psi[i] = np;
    red[0] += np * np;
  }
  block_reduce_store<1>(red, partials);
}

// psi = |psi| * sqrt(n4/(sum psi^2 dxyz)) [, max(psimin,.)], den = psi^2
__global__ void __launch_bounds__(256) k_renorm(long long n, double* psi, double* den, const double* scal,
                                                double n4, double dxyz, double psimin) {
  const double cnorm = sqrt(n4 / (scal[S_SUMSQ] * dxyz));
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double v = fabs(psi[i]) * cnorm;
    if (psimin > 0.0 && v < psimin) v = psimin;
    psi[i] = v;
    den[i] = v * v;
  }
}

// ---- steepest descent (He_statics/evolo.f90:5-96, ironing.f90:4-33)
// g = a1 (hpsi - mu psi), a1 = deltat/(1 - deltat (hmean - mu)) with mu := hmean
__global__ void __launch_bounds__(256) k_evolo_g(long long n, const double* hpsi, const double* psi,
                                                 const double* scal, double deltat, double* g) {
  const double hm = scal[S_MU];
  const double a1 = deltat / (1.0 - deltat * (hm - hm));
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    g[i] = a1 * (hpsi[i] - hm * psi[i]);
}
// one smoothing: interior points = (six neighbours)/12 + centre/2; boundary copied
__global__ void __launch_bounds__(256) k_ironing(Dims d, const double* in, double* out) {
#pragma clang fp contract(off)
  const long long sy = d.nx, sz = (long long)d.nx * d.ny;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < d.n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % d.nx), iy = (int)((i / d.nx) % d.ny), iz = (int)(i / sz);
    if (ix == 0 || iy == 0 || iz == 0 || ix == d.nx - 1 || iy == d.ny - 1 || iz == d.nz - 1) {
      out[i] = in[i];
    } else {
      double t = in[i - 1] + in[i + 1];
      t = t + in[i - sy];
      t = t + in[i + sy];
      t = t + in[i - sz];
      t = t + in[i + sz];
      out[i] = 8.333333333333333e-2 * t + 0.5 * in[i];
    }
  }
}
// psinw = psi - g + vdt1 (psi - psi1) + vdt2 (psi1 - psi2); shift the history; sum psinw^2
__global__ void __launch_bounds__(256) k_evolo_new(long long n, double* psi, double* psi1, double* psi2,
                                                   const double* g, double vdt1, double vdt2, double* partials) {
  double red[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double p = psi[i], p1 = psi1[i], p2 = psi2[i];
    const double nw = p - g[i] + vdt1 * (p - p1) + vdt2 * (p1 - p2);
    psi2[i] = p1;
    psi1[i] = p;
    psi[i] = nw;
    red[0] += nw * nw;
  }
  block_reduce_store<1>(red, partials);
}

// STEPPCR predictor + modifier (steppcr.f90:40-61), sum(psi^2)
__global__ void __launch_bounds__(256) k_pcr_pred(long long n, double* psi, const double* hpsi, double* pc,
                                                  double* timec, double* po3, const double* ho1, double* ho2,
                                                  const double* scal, double deltat, double dtmin, double sdmu,
                                                  double dmumax, double* partials) {
  double red[1] = {0.0};
  const double hmean = scal[S_MU];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double p = psi[i], hp = hpsi[i];
    const double tc = deltat / (1.0 + exp((fabs(hp / p) - dmumax) / sdmu)) + dtmin;
    timec[i] = tc;
    const double s4 = -(hp - hmean * p);
    const double s1 = po3[i] + C4O3 * tc * (2.0 * s4 - ho1[i] + 2.0 * ho2[i]);
    po3[i] = p;
    const double np = s1 - C112 * pc[i];
    psi[i] = np;
    pc[i] = s1;
    ho2[i] = s4;
    red[0] += np * np;
  }
  block_reduce_store<1>(red, partials);
}

// STEPPCR corrector + final (steppcr.f90:74-94): sum(psi^2), sum|pc|
__global__ void __launch_bounds__(256) k_pcr_corr(long long n, double* psi, const double* hpsi, double* pc,
                                                  const double* timec, const double* po3, const double* po2,
                                                  const double* ho1, const double* ho2, const double* scal,
                                                  double* partials) {
  double red[2] = {0.0, 0.0};
  const double mu = scal[S_MU];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double p = psi[i];
    const double s4 = -(hpsi[i] - mu * p);
    const double s5 = 0.125 * (9.0 * po3[i] - po2[i] + 3.0 * timec[i] * (s4 + 2.0 * ho1[i] - ho2[i]));
    const double pcv = pc[i] - s5;
    pc[i] = pcv;
    const double np = s5 + C9 * pcv;
    psi[i] = np;
    red[0] += np * np;
    red[1] += fabs(pcv);
  }
  block_reduce_store<2>(red, partials);
}

// ---- log-psi steppers (He_statics/steprkrl.f90, steppcrl.f90) --------------------
static constexpr double UNDERPSIL = -700.0;  // modules.f90:298

// out = acc + v (SQ: acc + v^2): the six derivative terms of H_l summed in the
// reference's order, (((d2x + d2y) + d2z) + d1x^2) + d1y^2 + d1z^2
template <bool SQ>
struct OutAccum {
  static constexpr int NR = 0;
  double* partials;
  double* out;
  const double* acc;
  __device__ __forceinline__ void store(size_t i, int, int, int, double v, double*) const {
    const double t = SQ ? v * v : v;
    out[i] = acc ? acc[i] + t : t;
  }
};

// hpsi = -g h2o2m4 + pot4 (steprkrl.f90:56), sum(den*hpsi)
__global__ void __launch_bounds__(256) k_hpsil(long long n, const double* g, const double* pot4, const double* den,
                                               double* hpsi, double h2o2m4, double* partials) {
  double red[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double hp = -g[i] * h2o2m4 + pot4[i];
    hpsi[i] = hp;
    red[0] += den[i] * hp;
  }
  block_reduce_store<1>(red, partials);
}

// psil = max(psil, underpsil); psi = exp(psil); accumulates sum(psi^2)
__device__ __forceinline__ double psil_exp(double* psil, double* psi, long long i, double v) {
  if (v < UNDERPSIL) v = UNDERPSIL;
  psil[i] = v;
  const double e = exp(v);
  psi[i] = e;
  return e * e;
}

// one RK-Gill stage of STEPRKRL after the mean (steprkrl.f90:61-77)
__global__ void __launch_bounds__(256) k_rkrl_stage(long long n, double* psil, double* psi, const double* hpsi,
                                                    double* q, double* po0, double* po1, double* po2, double* ho0,
                                                    double* ho1, const double* scal, double a, double b, double cc,
                                                    double deltat, int first, double* partials) {
  double red[1] = {0.0};
  const double mu = scal[S_MU];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double pl = psil[i];
    const double s4 = -(hpsi[i] - mu);
    const double s1 = a * (s4 - b * q[i]);
    q[i] = q[i] + 3.0 * s1 - cc * s4;
    if (first) {
      ho1[i] = ho0[i];
      ho0[i] = s4;
      po2[i] = po1[i];
      po1[i] = po0[i];
      po0[i] = pl;
    }
    red[0] += psil_exp(psil, psi, i, pl + deltat * s1);
  }
  block_reduce_store<1>(red, partials);
}

// cnorm = sqrt(n4/(sum psi^2 dxyz)); psil += log(cnorm); psi = |psi| cnorm; den = psi^2
// (steprkrl.f90:78-81)
__global__ void k_renorm_log(long long n, double* psil, double* psi, double* den, const double* scal, double n4,
                             double dxyz) {
  const double cnorm = sqrt(n4 / (scal[S_SUMSQ] * dxyz));
  const double lc = log(cnorm);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    psil[i] = psil[i] + lc;
    const double v = fabs(psi[i]) * cnorm;
    psi[i] = v;
    den[i] = v * v;
  }
}

// STEPPCRL predictor + modifier (steppcrl.f90:46-59), sum(psi^2)
__global__ void __launch_bounds__(256) k_pcrl_pred(long long n, double* psil, double* psi, const double* hpsi,
                                                   double* pc, double* po3, const double* ho1, double* ho2,
                                                   const double* scal, double deltat, double* partials) {
  double red[1] = {0.0};
  const double hmean = scal[S_MU];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double s4 = -(hpsi[i] - hmean);
    const double s1 = po3[i] + C4O3 * deltat * (2.0 * s4 - ho1[i] + 2.0 * ho2[i]);
    po3[i] = psil[i];
    const double np = s1 - C112 * pc[i];
    pc[i] = s1;
    ho2[i] = s4;
    red[0] += psil_exp(psil, psi, i, np);
  }
  block_reduce_store<1>(red, partials);
}

// STEPPCRL corrector + final (steppcrl.f90:86-98): sum(psi^2), sum|pc|
__global__ void __launch_bounds__(256) k_pcrl_corr(long long n, double* psil, double* psi, const double* hpsi,
                                                   double* pc, const double* po3, const double* po2,
                                                   const double* ho1, const double* ho2, const double* scal,
                                                   double deltat, double* partials) {
  double red[2] = {0.0, 0.0};
  const double mu = scal[S_MU];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double s4 = -(hpsi[i] - mu);
    const double s5 = 0.125 * (9.0 * po3[i] - po2[i] + 3.0 * deltat * (s4 + 2.0 * ho1[i] - ho2[i]));
    const double pcv = pc[i] - s5;
    pc[i] = pcv;
    red[0] += psil_exp(psil, psi, i, s5 + C9 * pcv);
    red[1] += fabs(pcv);
  }
  block_reduce_store<2>(red, partials);
}

// hpsix = hpsix - epsx psix with the previous eigenvalue (evolox.f90:50-52)
__global__ void k_evolox_g(long long n, double* hpsix, const double* psix, double epsx) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    hpsix[i] = hpsix[i] - epsx * psix[i];
}
// psixnw = psix - deltatx g + vdt1 (psix - psi1x) + vdt2 (psi1x - psi2x), history
// shift, max(psimin, .) (evolox.f90:63-77); sum psixnw^2
__global__ void __launch_bounds__(256) k_evolox_new(long long n, double* psix, double* psi1x, double* psi2x,
                                                    const double* g, double deltatx, double vdt1, double vdt2,
                                                    double psimin, double* partials) {
  double red[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double p = psix[i], p1 = psi1x[i], p2 = psi2x[i];
    double nw = p - deltatx * g[i] + vdt1 * (p - p1) + vdt2 * (p1 - p2);
    psi2x[i] = p1;
    psi1x[i] = p;
    nw = (psimin > nw) ? psimin : nw;
    psix[i] = nw;
    red[0] += nw * nw;
  }
  block_reduce_store<1>(red, partials);
}
// psix = psixnw sqrt(1/(sum psixnw^2 dxyz)) (evolox.f90:78-83)
__global__ void k_norm1(long long n, double* psix, const double* scal, double dxyz) {
  const double cnorm = sqrt(1.0 / (scal[S_SUMSQ] * dxyz));
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    psix[i] = psix[i] * cnorm;
}

__global__ void k_log_abs(long long n, const double* psi, double* psil) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    psil[i] = log(fabs(psi[i]));
}

__global__ void k_copy_scal(double* s, int from, int to) { s[to] = s[from]; }

// statics start: den *= n4/(sum den dxyz), psi = sqrt(den) (BCN4HeDFT.f90:730-734)
__global__ void k_sum1(long long n, const double* a, double* partials) {
  double red[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    red[0] += a[i];
  block_reduce_store<1>(red, partials);
}
__global__ void k_statics_psi0(long long n, double* den, double* psi, const double* scal, double n4, double dxyz) {
  const double cnorm = n4 / (scal[S_SUMSQ] * dxyz);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double v = cnorm * den[i];
    den[i] = v;
    psi[i] = sqrt(v);
  }
}

// ---- statics -------------------------------------------------------------------------------------
static int hpsi_statics(hedft_ctx* c, double n4) {
  HEDFT_TRY(lap_xy(c, c->psi, 0));
  OutHpsiR o;
  o.partials = gen_partials(c);
  o.lapxy = rz(c->lap);
  o.psi = rz(c->psi);
  o.pot4 = c->pot4;
  o.hpsi = c->hpsi_r;
  o.h2o2m4 = c->p.h2o2m4;
  launch_pencil<double, 2>(c->stream, c->d, c->coef[1][2], InPlain<double>{rz(c->psi)}, o);
  finalize<1>(c, gen_partials(c), pencil_geom<double>(c->d, 2).nblocks, S_MU, c->dxyz / n4);
  return HEDFT_OK;
}

static void renorm(hedft_ctx* c, double n4) {
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  hipLaunchKernelGGL(k_renorm, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->den, c->scal, n4,
                     c->dxyz, c->p.psimin);
}

// renormalisation without the psimin floor (evolo.f90:84-94)
static void renorm_nomin(hedft_ctx* c, double n4) {
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  hipLaunchKernelGGL(k_renorm, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->den, c->scal, n4,
                     c->dxyz, -1.0);
}

// STEPRKR (He_statics/steprkr.f90:1-96)
static int steprkr_impl(hedft_ctx* c, double dt, double n4, double* hmean) {
  for (int jrun = 0; jrun < 4; ++jrun) {
    HEDFT_TRY(hpsi_statics(c, n4));
    if (jrun == 0)
      hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, c->stream, c->scal, S_MU, S_HMEAN);
    hipLaunchKernelGGL(k_rkr_stage, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->hpsi_r,
                       rz(c->q), c->timec_r, rz(c->psiold[0]), rz(c->psiold[1]), rz(c->psiold[2]),
                       rz(c->hpsiold[0]), rz(c->hpsiold[1]), c->scal, RK_A[jrun], RK_B[jrun], RK_C[jrun], dt,
                       c->p.dtmin, c->p.sdmu, c->p.dmumax, c->p.psimin, (int)(jrun == 0), gen_partials(c));
    // psimin clamp is applied before the norm in steprkr (steprkr.f90:73-78)
    finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
    hipLaunchKernelGGL(k_renorm, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->den, c->scal,
                       n4, c->dxyz, -1.0);
    if (jrun <= 2) HEDFT_TRY(poten_statics(c));
  }
  for (int k = 0; k < 3; ++k) c->ioldp[k] = k;
  for (int k = 0; k < 2; ++k) c->ioldh[k] = k;
  c->rk_done++;
  (void)hmean;
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

// STEPPCR (He_statics/steppcr.f90:1-108)
static int steppcr_impl(hedft_ctx* c, double dt, double n4) {
  HEDFT_TRY(hpsi_statics(c, n4));
  hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, c->stream, c->scal, S_MU, S_HMEAN);
  hipLaunchKernelGGL(k_pcr_pred, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->hpsi_r,
                     rz(c->pc), c->timec_r, rz(c->psiold[c->ioldp[2]]), rz(c->hpsiold[c->ioldh[0]]),
                     rz(c->hpsiold[c->ioldh[1]]), c->scal, dt, c->p.dtmin, c->p.sdmu, c->p.dmumax, gen_partials(c));
  renorm(c, n4);
  std::swap(c->ioldh[0], c->ioldh[1]);
  HEDFT_TRY(poten_statics(c));
  HEDFT_TRY(hpsi_statics(c, n4));
  hipLaunchKernelGGL(k_pcr_corr, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->hpsi_r,
                     rz(c->pc), c->timec_r, rz(c->psiold[c->ioldp[2]]), rz(c->psiold[c->ioldp[1]]),
                     rz(c->hpsiold[c->ioldh[0]]), rz(c->hpsiold[c->ioldh[1]]), c->scal, gen_partials(c));
  hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(256), 0, c->stream, gen_partials(c), EW_BLOCKS,
                     c->scal + S_SUMSQ, 1.0);  // S_SUMSQ = sum psi^2, S_ERRPC = sum|pc|
  hipLaunchKernelGGL(k_renorm, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->den, c->scal, n4,
                     c->dxyz, c->p.psimin);
  const int t = c->ioldp[2];
  c->ioldp[2] = c->ioldp[1];
  c->ioldp[1] = c->ioldp[0];
  c->ioldp[0] = t;
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

// one iteration of the statics driver loop (BCN4HeDFT.f90:1113-1160): the
// stepper, then the driver's refresh (derden, fden, dencg, denalf) and poten
int statics_rk_iter(hedft_ctx* c, double dt, double n4) {
  HEDFT_TRY(steprkr_impl(c, dt, n4, nullptr));
  HEDFT_TRY(statics_refresh(c));
  return poten_statics(c);
}
int statics_pc_iter(hedft_ctx* c, double dt, double n4) {
  HEDFT_TRY(steppcr_impl(c, dt, n4));
  HEDFT_TRY(statics_refresh(c));
  return poten_statics(c);
}

// H_l psil (steprkrl.f90:44-57): six derivative passes of psil with one-sided
// rows, accumulated in work1, then hpsi and sum(den*hpsi)*dxyz/n4 into S_MU
static int hpsi_log(hedft_ctx* c, double n4) {
  double* g = c->work1;
  const double* pl = c->psil;
  InPlain<double> ip{pl};
  launch_pencil<double, 0>(c->stream, c->d, c->coef0[1][0], ip, OutAccum<false>{nullptr, g, nullptr});
  launch_pencil<double, 1>(c->stream, c->d, c->coef0[1][1], ip, OutAccum<false>{nullptr, g, g});
  launch_pencil<double, 2>(c->stream, c->d, c->coef0[1][2], ip, OutAccum<false>{nullptr, g, g});
  launch_pencil<double, 0>(c->stream, c->d, c->coef0[0][0], ip, OutAccum<true>{nullptr, g, g});
  launch_pencil<double, 1>(c->stream, c->d, c->coef0[0][1], ip, OutAccum<true>{nullptr, g, g});
  launch_pencil<double, 2>(c->stream, c->d, c->coef0[0][2], ip, OutAccum<true>{nullptr, g, g});
  hipLaunchKernelGGL(k_hpsil, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, g, c->pot4, c->den, c->hpsi_r,
                     c->p.h2o2m4, gen_partials(c));
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_MU, c->dxyz / n4);
  return HEDFT_OK;
}

static void renorm_log(hedft_ctx* c, double n4) {
  hipLaunchKernelGGL(k_renorm_log, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->psil, rz(c->psi), c->den,
                     c->scal, n4, c->dxyz);
}

// STEPRKRL (He_statics/steprkrl.f90:1-98)
static int steprkrl_impl(hedft_ctx* c, double dt, double n4) {
  for (int jrun = 0; jrun < 4; ++jrun) {
    HEDFT_TRY(hpsi_log(c, n4));
    if (jrun == 0) hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, c->stream, c->scal, S_MU, S_HMEAN);
    hipLaunchKernelGGL(k_rkrl_stage, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->psil, rz(c->psi),
                       c->hpsi_r, rz(c->q), rz(c->psiold[0]), rz(c->psiold[1]), rz(c->psiold[2]),
                       rz(c->hpsiold[0]), rz(c->hpsiold[1]), c->scal, RK_A[jrun], RK_B[jrun], RK_C[jrun], dt,
                       (int)(jrun == 0), gen_partials(c));
    finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
    renorm_log(c, n4);
    if (jrun <= 2) HEDFT_TRY(poten_statics(c));
  }
  for (int k = 0; k < 3; ++k) c->ioldp[k] = k;
  for (int k = 0; k < 2; ++k) c->ioldh[k] = k;
  c->rk_done++;
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

// STEPPCRL (He_statics/steppcrl.f90:1-116)
static int steppcrl_impl(hedft_ctx* c, double dt, double n4) {
  HEDFT_TRY(hpsi_log(c, n4));
  hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, c->stream, c->scal, S_MU, S_HMEAN);
  hipLaunchKernelGGL(k_pcrl_pred, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->psil, rz(c->psi),
                     c->hpsi_r, rz(c->pc), rz(c->psiold[c->ioldp[2]]), rz(c->hpsiold[c->ioldh[0]]),
                     rz(c->hpsiold[c->ioldh[1]]), c->scal, dt, gen_partials(c));
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  renorm_log(c, n4);
  std::swap(c->ioldh[0], c->ioldh[1]);
  HEDFT_TRY(poten_statics(c));
  HEDFT_TRY(hpsi_log(c, n4));
  hipLaunchKernelGGL(k_pcrl_corr, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->psil, rz(c->psi),
                     c->hpsi_r, rz(c->pc), rz(c->psiold[c->ioldp[2]]), rz(c->psiold[c->ioldp[1]]),
                     rz(c->hpsiold[c->ioldh[0]]), rz(c->hpsiold[c->ioldh[1]]), c->scal, dt, gen_partials(c));
  hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(256), 0, c->stream, gen_partials(c), EW_BLOCKS,
                     c->scal + S_SUMSQ, 1.0);  // S_SUMSQ = sum psi^2, S_ERRPC = sum|pc|
  renorm_log(c, n4);
  const int t = c->ioldp[2];
  c->ioldp[2] = c->ioldp[1];
  c->ioldp[1] = c->ioldp[0];
  c->ioldp[0] = t;
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

}  // namespace hedft

extern "C" {

int hedft_steprkr(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  HEDFT_TRY(steprkr_impl(c, deltat, n4, nullptr));
  double h;
  HEDFT_TRY(read_scal(c, S_HMEAN, 1, &h));
  if (mu4err) *mu4err = std::fabs(1.0 - (mu4 ? *mu4 : 0.0) / h);
  if (mu4) *mu4 = h;
  return done(c);
}

int hedft_steppcr(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4, double* errpc) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  if (!c->rk_done) return HEDFT_ESTATE;
  HEDFT_TRY(steppcr_impl(c, deltat, n4));
  double s[4];
  HEDFT_TRY(read_scal(c, 0, 4, s));
  if (mu4err) *mu4err = std::fabs(1.0 - (mu4 ? *mu4 : 0.0) / s[S_HMEAN]);
  if (mu4) *mu4 = s[S_HMEAN];
  if (errpc) *errpc = (9.0 / 121.0) * s[S_ERRPC] / (double)c->d.n;
  return done(c);
}

int hedft_statics_log_start(hedft_ctx* c) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  hipLaunchKernelGGL(k_log_abs, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->psil);
  return done(c);
}

int hedft_steprkrl(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  HEDFT_TRY(steprkrl_impl(c, deltat, n4));
  double h;
  HEDFT_TRY(read_scal(c, S_HMEAN, 1, &h));
  if (mu4err) *mu4err = std::fabs(1.0 - (mu4 ? *mu4 : 0.0) / h);
  if (mu4) *mu4 = h;
  return done(c);
}

int hedft_steppcrl(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4, double* errpc) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  if (!c->rk_done) return HEDFT_ESTATE;
  HEDFT_TRY(steppcrl_impl(c, deltat, n4));
  double s[4];
  HEDFT_TRY(read_scal(c, 0, 4, s));
  if (mu4err) *mu4err = std::fabs(1.0 - (mu4 ? *mu4 : 0.0) / s[S_HMEAN]);
  if (mu4) *mu4 = s[S_HMEAN];
  if (errpc) *errpc = (9.0 / 121.0) * s[S_ERRPC] / (double)c->d.n;
  return done(c);
}

int hedft_statics_init(hedft_ctx* c, double n4, double* mu4) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  hipLaunchKernelGGL(k_sum1, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->den, gen_partials(c));
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  hipLaunchKernelGGL(k_statics_psi0, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->den, rz(c->psi), c->scal,
                     n4, c->dxyz);
  if (c->quantum) {  // psi1x = psi2x = psix (BCN4HeDFT.f90:781-785)
    const size_t b = sizeof(double) * c->d.n;
    HEDFT_CHECK(hipMemcpyAsync(c->qi[0].psixold[0], c->qi[0].psix, b, hipMemcpyDeviceToDevice, c->stream));
    HEDFT_CHECK(hipMemcpyAsync(c->qi[0].psixold[1], c->qi[0].psix, b, hipMemcpyDeviceToDevice, c->stream));
  }
  HEDFT_TRY(statics_refresh(c));
  HEDFT_TRY(poten_statics(c));
  HEDFT_TRY(hpsi_statics(c, n4));
  // psi1 = psi2 = psi for the steepest-descent method (BCN4HeDFT.f90:861-862)
  HEDFT_CHECK(hipMemcpyAsync(c->psi1s, c->psi, sizeof(double) * c->d.n, hipMemcpyDeviceToDevice, c->stream));
  HEDFT_CHECK(hipMemcpyAsync(c->psi2s, c->psi, sizeof(double) * c->d.n, hipMemcpyDeviceToDevice, c->stream));
  double m;
  HEDFT_TRY(read_scal(c, S_MU, 1, &m));
  if (mu4) *mu4 = m;
  c->rk_done = 0;
  return done(c);
}

int hedft_evolo(hedft_ctx* c, double deltat, double vdt1, double vdt2, int32_t iron, double n4, double* mu4,
                double* mu4err) {
  if (!c || c->mode != HEDFT_MODE_STATICS || iron < 0 || !mu4) return HEDFT_EINVAL;
  HEDFT_TRY(hpsi_statics(c, n4));
  double* g = c->work1;
  hipLaunchKernelGGL(k_evolo_g, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->hpsi_r, rz(c->psi), c->scal,
                     deltat, g);
  for (int k = 0; k < iron; ++k) {  // ping-pong between work1 and work2
    double* dst = (g == c->work1) ? c->work2 : c->work1;
    hipLaunchKernelGGL(k_ironing, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, g, dst);
    g = dst;
  }
  hipLaunchKernelGGL(k_evolo_new, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psi), c->psi1s, c->psi2s,
                     g, vdt1, vdt2, gen_partials(c));
  renorm_nomin(c, n4);
  double hmean;
  HEDFT_TRY(read_scal(c, S_MU, 1, &hmean));
  if (mu4err) *mu4err = std::fabs(1.0 - *mu4 / hmean);
  *mu4 = hmean;
  return done(c);
}

// evolox (He_statics/evolox.f90:5-87)
int hedft_evolox(hedft_ctx* c, double deltatx, double vdt1, double vdt2, int32_t ironx, double* epsx,
                 double* epsxerr) {
  if (!c || c->mode != HEDFT_MODE_STATICS || !c->quantum || ironx < 0 || !epsx) return HEDFT_EINVAL;
  // hpsix = -(lap psix) h2o2mx + upotx psix, sum(psix hpsix) dxyz (evolox.f90:39-48)
  HEDFT_TRY(lap_xy(c, c->qi[0].psix, 0));
  OutHpsiR o;
  o.partials = gen_partials(c);
  o.lapxy = rz(c->lap);
  o.psi = rz(c->qi[0].psix);
  o.pot4 = c->qi[0].upotx;
  o.hpsi = c->hpsi_r;
  o.h2o2m4 = c->qi[0].h2o2mimp;
  launch_pencil<double, 2>(c->stream, c->d, c->coef[1][2], InPlain<double>{rz(c->qi[0].psix)}, o);
  finalize<1>(c, gen_partials(c), pencil_geom<double>(c->d, 2).nblocks, S_MU, c->dxyz);
  const double eold = *epsx;
  hipLaunchKernelGGL(k_evolox_g, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->hpsi_r, rz(c->qi[0].psix), eold);
  double* g = c->hpsi_r;
  for (int k = 0; k < ironx; ++k) {  // ping-pong through work1 / work2
    double* dst = (g == c->work1) ? c->work2 : c->work1;
    hipLaunchKernelGGL(k_ironing, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, g, dst);
    g = dst;
  }
  hipLaunchKernelGGL(k_evolox_new, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->qi[0].psix),
                     rz(c->qi[0].psixold[0]), rz(c->qi[0].psixold[1]), g, deltatx, vdt1, vdt2, c->p.psimin, gen_partials(c));
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  hipLaunchKernelGGL(k_norm1, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->qi[0].psix), c->scal, c->dxyz);
  double hmean;
  HEDFT_TRY(read_scal(c, S_MU, 1, &hmean));
  if (epsxerr) *epsxerr = std::fabs(1.0 - eold / hmean);
  *epsx = hmean;
  return done(c);
}

// the driver's part after the steps (BCN4HeDFT.f90:1125-1160): derden, fden
// [, fdenx], dencg, denalf, poten
int hedft_statics_refresh(hedft_ctx* c) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  HEDFT_TRY(statics_refresh(c));
  HEDFT_TRY(poten_statics(c));
  return done(c);
}

int hedft_statics_iteration_sd(hedft_ctx* c, double deltat, double vdt1, double vdt2, int32_t iron, double n4,
                               double* mu4, double* mu4err) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  HEDFT_TRY(hedft_evolo(c, deltat, vdt1, vdt2, iron, n4, mu4, mu4err));
  HEDFT_TRY(statics_refresh(c));
  HEDFT_TRY(poten_statics(c));
  return done(c);
}

int hedft_statics_iteration_log(hedft_ctx* c, double deltat, int32_t iterrk, double n4, double* mu4,
                                double* mu4err, double* errpc) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  if (iterrk <= 4) {
    HEDFT_TRY(hedft_steprkrl(c, deltat, mu4, mu4err, n4));
    if (errpc) *errpc = 0.0;
  } else {
    HEDFT_TRY(hedft_steppcrl(c, deltat, mu4, mu4err, n4, errpc));
  }
  HEDFT_TRY(statics_refresh(c));
  HEDFT_TRY(poten_statics(c));
  return done(c);
}

int hedft_statics_iteration(hedft_ctx* c, double deltat, int32_t iterrk, double n4, double* mu4, double* mu4err) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  double errpc;
  if (iterrk <= 4) HEDFT_TRY(hedft_steprkr(c, deltat, mu4, mu4err, n4));
  else HEDFT_TRY(hedft_steppcr(c, deltat, mu4, mu4err, n4, &errpc));
  HEDFT_TRY(statics_refresh(c));
  HEDFT_TRY(poten_statics(c));
  return done(c);
}

}  // extern "C"
