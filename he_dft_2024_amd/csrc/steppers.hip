// steppers.hip — the real-time integrators: the fused H psi epilogues of the
// RK-Gill stage, Hamming predictor and corrector (STEPRK / STEPPC of
// Dynamics_TDDFT and Dynamics_quantum_impurities), the classical impurities'
// Newton updates, and the small kernels a step leaves for the next mean field.
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

// ---- fused z-pass epilogues of the TDDFT integrators -----------------------------
// k = timec*((lap)*h2o2m4 - (pot4-mu4)*psi) - i*uimp*psi   (steprk.f90:67,
// steppc.f90:51,118; mu4 and no uimp in Dynamics_quantum_impurities/steprk.f90:68,
// steppc.f90:46,105; mu4 = 0 for classical impurities)
// ps is psi at the point, taken from the stencil's LDS tile (not re-read)
template <typename R>
struct HpsiC {
  const double2* lapxy;
  const R* pot4;
  Lam lam;
  ImpDev im;
  double h2o2m4, mu4;
  // with one impurity its table lookup is part of the operands (loaded
  // ahead by the 2.5-D kernel); with several it is evaluated in kv
  __device__ __forceinline__ HOps ops(size_t i, int ix, int iy, int iz) const {
    HOps o{(double)pot4[i], lam.tx[ix], lam.ty[iy], lam.tz[iz], 0.0, 0.0, 0.0};
    if (im.nimp == 1) {
      double xx, yy, zz;
      imp_xyz(im, ix, iy, iz, xx, yy, zz);
      int ir;
      imp_node_k(im, 0, xx, yy, zz, false, ir, o.rm);
      o.t0 = im.potion[ir];
      o.t1 = im.potion[ir + 1];
    }
    return o;
  }
  __device__ __forceinline__ double2 k(size_t i, int ix, int iy, int iz, double2 dz, double2 ps) const {
    return kv(i, ix, iy, iz, dz, ps, ops(i, ix, iy, iz));
  }
  // o: operands at the point; lapxy null: dz is the whole Laplacian (the
  // 2.5-D stepper kernel)
  __device__ __forceinline__ double2 kv(size_t i, int ix, int iy, int iz, double2 dz, double2 ps,
                                        const HOps& o) const {
    double2 lap = dz;
    if (lapxy) {
      const double2 l0 = lapxy[i];
      lap = make_double2(l0.x + dz.x, l0.y + dz.y);
    }
    const double v = o.p - mu4;
    const double2 t = make_double2(lap.x * h2o2m4 - v * ps.x, lap.y * h2o2m4 - v * ps.y);
    const double lr = lam.of(o.ax, o.ay, o.az);
    // (lr + i) * t
    const double2 tt = make_double2(lr * t.x - t.y, lr * t.y + t.x);
    double u = 0.0;
    if (im.nimp == 1) u += imp_lerp(o.t0, o.t1, o.rm);
    else if (im.nimp > 1) u = imp_u(im, ix, iy, iz);
    // - i u psi = (u psi.y, -u psi.x)
    return make_double2(tt.x + u * ps.y, tt.y - u * ps.x);
  }
};

// quantum impurity: k = timec*((lap psix)*h2o2mimp - upotx*psix)
// (Dynamics_quantum_impurities/steprk.f90:94, steppc.f90:77,127)
template <typename R>
struct HpsiX {
  const double2* lapxy;
  const R* upot;
  Lam lam;
  double h2o2m;
  __device__ __forceinline__ HOps ops(size_t i, int ix, int iy, int iz) const {
    return HOps{(double)upot[i], lam.tx[ix], lam.ty[iy], lam.tz[iz], 0.0, 0.0, 0.0};
  }
  __device__ __forceinline__ double2 k(size_t i, int ix, int iy, int iz, double2 dz, double2 ps) const {
    return kv(i, ix, iy, iz, dz, ps, ops(i, ix, iy, iz));
  }
  __device__ __forceinline__ double2 kv(size_t i, int ix, int iy, int iz, double2 dz, double2 ps,
                                        const HOps& o) const {
    double2 lap = dz;
    if (lapxy) {
      const double2 l0 = lapxy[i];
      lap = make_double2(l0.x + dz.x, l0.y + dz.y);
    }
    const double v = o.p;
    const double2 t = make_double2(lap.x * h2o2m - v * ps.x, lap.y * h2o2m - v * ps.y);
    const double lr = lam.of(o.ax, o.ay, o.az);
    return make_double2(lr * t.x - t.y, lr * t.y + t.x);
  }
};

// Outputs per sliding window in the fused stepper epilogues (each output
// touches 6-8 complex arrays, so fewer outputs are kept in flight).
#ifndef HEDFT_KB_STEP
#define HEDFT_KB_STEP 1
#endif

// H: HpsiC (helium) or HpsiX (quantum impurity); den (|psi|^2, R storage) may be null
template <typename R, typename H = HpsiC<R>>
struct OutRK {  // one RK-Gill stage (steprk.f90:67-79)
  static constexpr int NR = 0;
  static constexpr int KB = HEDFT_KB_STEP;
  static constexpr bool CENTRE = true;  // store() receives psi at the point from the LDS tile
  double* partials;
  H h;
  double2 *psi, *q, *po0, *po1, *po2, *ho0, *ho1;
  R* den;
  double a, b, cc, dt;
  int first;
  // the operands read at the point, loaded ahead of the arithmetic by the
  // 2.5-D kernel (store() loads and applies them at once)
  struct Ops {
    double2 qq, h0, p1, p0;
    HOps v;
  };
  // (the history slots are read on every stage, used on the first)
  __device__ __forceinline__ Ops load(size_t i, int ix, int iy, int iz) const {
    return Ops{q[i], ho0[i], po1[i], po0[i], h.ops(i, ix, iy, iz)};
  }
  __device__ __forceinline__ void store(size_t i, int ix, int iy, int iz, double2 dz, double2 ps, double* red) const {
    apply(i, ix, iy, iz, dz, ps, load(i, ix, iy, iz), red);
  }
  __device__ __forceinline__ void apply(size_t i, int ix, int iy, int iz, double2 dz, double2 ps, const Ops& o,
                                        double*) const {
    const double2 k4 = h.kv(i, ix, iy, iz, dz, ps, o.v);
    const double2 qq = o.qq;
    const double2 s1 = make_double2(a * (k4.x - b * qq.x), a * (k4.y - b * qq.y));
    q[i] = make_double2(qq.x + 3.0 * s1.x - cc * k4.x, qq.y + 3.0 * s1.y - cc * k4.y);
    if (first) {
      ho1[i] = o.h0;
      ho0[i] = k4;
      po2[i] = o.p1;
      po1[i] = o.p0;
      po0[i] = ps;
    }
    const double2 np = make_double2(ps.x + dt * s1.x, ps.y + dt * s1.y);
    psi[i] = np;
    if (den) den[i] = (R)(np.x * np.x + np.y * np.y);
  }
};

template <typename R, typename H = HpsiC<R>>
struct OutPred {  // predictor + modifier (steppc.f90:51-61)
  static constexpr int NR = 0;
  static constexpr int KB = HEDFT_KB_STEP;
  static constexpr bool CENTRE = true;  // store() receives psi at the point from the LDS tile
  double* partials;
  H h;
  double2 *psi, *pc, *po3, *ho1, *ho2;
  R* den;
  double dt;
  struct Ops {
    double2 o3, h1, h2, p;
    HOps v;
  };
  __device__ __forceinline__ Ops load(size_t i, int ix, int iy, int iz) const {
    return Ops{po3[i], ho1[i], ho2[i], pc[i], h.ops(i, ix, iy, iz)};
  }
  __device__ __forceinline__ void store(size_t i, int ix, int iy, int iz, double2 dz, double2 ps, double* red) const {
    apply(i, ix, iy, iz, dz, ps, load(i, ix, iy, iz), red);
  }
  __device__ __forceinline__ void apply(size_t i, int ix, int iy, int iz, double2 dz, double2 ps, const Ops& o,
                                        double*) const {
    const double2 k4 = h.kv(i, ix, iy, iz, dz, ps, o.v);
    const double2 o3 = o.o3, h1 = o.h1, h2 = o.h2;
    const double f = C4O3 * dt;
    const double2 s1 = make_double2(o3.x + f * (2.0 * k4.x - h1.x + 2.0 * h2.x),
                                    o3.y + f * (2.0 * k4.y - h1.y + 2.0 * h2.y));
    po3[i] = ps;
    const double2 p = o.p;
    const double2 np = make_double2(s1.x - C112 * p.x, s1.y - C112 * p.y);
    psi[i] = np;
    pc[i] = s1;
    ho2[i] = k4;
    if (den) den[i] = (R)(np.x * np.x + np.y * np.y);
  }
};

template <typename R, typename H = HpsiC<R>>
struct OutCorr {  // corrector + final value (steppc.f90:118-127), sum|pc|
  static constexpr int NR = 1;
  static constexpr int KB = HEDFT_KB_STEP;
  static constexpr bool CENTRE = true;  // store() receives psi at the point from the LDS tile
  double* partials;
  H h;
  double2 *psi, *pc, *po3, *po2, *ho1, *ho2;
  R* den;
  double dt;
  struct Ops {
    double2 o3, o2, h1, h2, p0;
    HOps v;
  };
  __device__ __forceinline__ Ops load(size_t i, int ix, int iy, int iz) const {
    return Ops{po3[i], po2[i], ho1[i], ho2[i], pc[i], h.ops(i, ix, iy, iz)};
  }
  __device__ __forceinline__ void store(size_t i, int ix, int iy, int iz, double2 dz, double2 ps, double* red) const {
    apply(i, ix, iy, iz, dz, ps, load(i, ix, iy, iz), red);
  }
  __device__ __forceinline__ void apply(size_t i, int ix, int iy, int iz, double2 dz, double2 ps, const Ops& o,
                                        double* red) const {
    const double2 k4 = h.kv(i, ix, iy, iz, dz, ps, o.v);
    const double2 o3 = o.o3, o2 = o.o2, h1 = o.h1, h2 = o.h2;
    const double f = 3.0 * dt;
    const double2 s5 = make_double2(0.125 * (9.0 * o3.x - o2.x + f * (k4.x + 2.0 * h1.x - h2.x)),
                                    0.125 * (9.0 * o3.y - o2.y + f * (k4.y + 2.0 * h1.y - h2.y)));
    const double2 p0 = o.p0;
    const double2 p = make_double2(p0.x - s5.x, p0.y - s5.y);
    pc[i] = p;
    const double2 np = make_double2(s5.x + C9 * p.x, s5.y + C9 * p.y);
    psi[i] = np;
    if (den) den[i] = (R)(np.x * np.x + np.y * np.y);
    red[0] += sqrt(p.x * p.x + p.y * p.y);
  }
};

// ---- impurity Newton updates (one lane per coordinate; N_imp*3 <= 12 numbers) --------
// Every coordinate's update reads and writes only its own entries of the state
// arrays, so lane i carries coordinate i through the reference's loops in
// their order; launched as one wavefront.
__global__ void k_imp_rk(double* imp, int nimp, int first, double a, double b, double cc, double dt) {
  const int m = 3 * nimp, i = threadIdx.x;
  if (i >= m) return;
  double *R = imp + IMP_R, *V = imp + IMP_V, *A = imp + IMP_A, *QR = imp + IMP_QR, *QV = imp + IMP_QV;
  double *RO = imp + IMP_ROLD, *VO = imp + IMP_VOLD, *AO = imp + IMP_AOLD;
  {  // steprk.f90:93-103
    const double st = a * (V[i] - b * QR[i]);
    QR[i] = QR[i] + 3.0 * st - cc * V[i];
    if (first) {
      RO[24 + i] = RO[12 + i];
      RO[12 + i] = RO[i];
      RO[i] = R[i];
    }
    R[i] = R[i] + dt * st;
  }
  {  // steprk.f90:108-117
    const double st = a * (A[i] - b * QV[i]);
    QV[i] = QV[i] + 3.0 * st - cc * A[i];
    if (first) {
      AO[12 + i] = AO[i];
      AO[i] = A[i];
      VO[24 + i] = VO[12 + i];
      VO[12 + i] = VO[i];
      VO[i] = V[i];
    }
    V[i] = V[i] + dt * st;
  }
}

// steppc.f90:79-99 (slot indices 0-based)
__global__ void k_imp_pred(double* imp, int nimp, double dt, int r3, int v1, int v2, int v3, int a1, int a2) {
  const int m = 3 * nimp, i = threadIdx.x;
  if (i >= m) return;
  double *R = imp + IMP_R, *V = imp + IMP_V, *A = imp + IMP_A, *PCR = imp + IMP_PCR, *PCV = imp + IMP_PCV;
  double *RO = imp + IMP_ROLD, *VO = imp + IMP_VOLD, *AO = imp + IMP_AOLD;
  const double sr = RO[12 * r3 + i] + C4O3 * dt * (2.0 * V[i] - VO[12 * v1 + i] + 2.0 * VO[12 * v2 + i]);
  RO[12 * r3 + i] = R[i];
  R[i] = sr - C112 * PCR[i];
  PCR[i] = sr;
  const double sv = VO[12 * v3 + i] + C4O3 * dt * (2.0 * A[i] - AO[12 * a1 + i] + 2.0 * AO[12 * a2 + i]);
  VO[12 * v3 + i] = sv - C112 * PCV[i];
  PCV[i] = sv;
  AO[12 * a2 + i] = A[i];
}

// steppc.f90:148-174; errimp, errvimp into scal (summed by lane 0 in coordinate order)
__global__ void k_imp_corr(double* imp, int nimp, double dt, int r2, int r3, int v1, int v2, int v3, int a1,
                           int a2, double* scal) {
  __shared__ double err[2][3 * HEDFT_MAX_IMP];
  const int m = 3 * nimp, i = threadIdx.x;
  if (i < m) {
    double *R = imp + IMP_R, *V = imp + IMP_V, *A = imp + IMP_A, *PCR = imp + IMP_PCR, *PCV = imp + IMP_PCV;
    double *RO = imp + IMP_ROLD, *VO = imp + IMP_VOLD, *AO = imp + IMP_AOLD;
    const double sr = 0.125 * (9.0 * RO[12 * r3 + i] - RO[12 * r2 + i] +
                               3.0 * dt * (VO[12 * v3 + i] + 2.0 * V[i] - VO[12 * v1 + i]));
    const double pr = PCR[i] - sr;
    PCR[i] = pr;
    R[i] = sr + C9 * pr;
    err[0][i] = fabs(C9 * pr);
    const double sv = 0.125 * (9.0 * V[i] - VO[12 * v2 + i] + 3.0 * dt * (A[i] + 2.0 * AO[12 * a1 + i] - AO[12 * a2 + i]));
    const double pv = PCV[i] - sv;
    PCV[i] = pv;
    VO[12 * v3 + i] = V[i];
    V[i] = sv + C9 * pv;
    err[1][i] = fabs(C9 * pv);
  }
  __syncthreads();
  if (i == 0) {
    double e1 = 0.0, e2 = 0.0;
    for (int k = 0; k < m; ++k) {
      e1 += err[0][k];
      e2 += err[1][k];
    }
    scal[S_ERRIMP] = e1 * 0.3333333333 / nimp;
    scal[S_ERRVIMP] = e2 * 0.3333333333 / nimp;
  }
}

__global__ void k_imp_aimp(double* imp, int nimp) {
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < nimp; ++k) imp[IMP_A + k + nimp * a] = imp[IMP_F + k + nimp * a] / imp[IMP_MASS + k];
}

// errHe reduction); also called directly where no mean field follows
void flush_side_errors(hedft_ctx* c) {
  if (c->side_pre & 4)
    finalize<1>(c, gen_partials(c), c->side_nblocks, S_ERRPC, 1.0);
  if (c->side_pre & 8)
    for (int k = 0; k < c->quantum; ++k) finalize<1>(c, x_partials(c, k), c->side_nblocks, S_ERRPCX + k, 1.0);
  c->side_pre &= ~(4 | 8);
}
void flush_side_impurity(hedft_ctx* c) {
  const int* a = c->side_args;
  if (c->side_pre & 1)
    hipLaunchKernelGGL(k_imp_pred, dim3(1), dim3(64), 0, c->stream, c->imp, c->nimp, c->side_dt, a[0], a[1], a[2],
                       a[3], a[4], a[5]);
  if (c->side_pre & 2)
    hipLaunchKernelGGL(k_imp_corr, dim3(1), dim3(64), 0, c->stream, c->imp, c->nimp, c->side_dt, a[0], a[1], a[2],
                       a[3], a[4], a[5], a[6], c->scal);
  c->side_pre &= ~(1 | 2);
}
void flush_side_pre(hedft_ctx* c) {
  flush_side_errors(c);
  flush_side_impurity(c);
}

// potx4 = F^-1[F(|psix|^2) uimp_k] of a quantum impurity
// (Dynamics_quantum_impurities/potenimp.f90:22-27), one forward, one
// ---- integrator drivers ------------------------------------------------------------------------
Lam make_lam(hedft_ctx* c) {
  Lam l;
  l.tx = c->lamtab;
  l.ty = c->lamtab + c->d.nx;
  l.tz = c->lamtab + c->d.nx + c->d.ny;
  l.lambdah = c->p.lambdah;
  return l;
}

template <typename R>
static HpsiC<R> make_hpsic(hedft_ctx* c) {
  HpsiC<R> h;
  h.lapxy = c->lap25_tz > 0 ? nullptr : (const double2*)c->lap;
  h.pot4 = fld<R>(c->pot4);
  h.lam = make_lam(c);
  h.im = make_impdev(c);
  h.h2o2m4 = c->p.h2o2m4;
  h.mu4 = c->quantum ? c->mu4 : 0.0;
  return h;
}

template <typename R>
static HpsiX<R> make_hpsix(hedft_ctx* c, int k) {
  HpsiX<R> h;
  h.lapxy = c->lap25_tz > 0 ? nullptr : (const double2*)c->qi[k].lapx;
  h.upot = fld<R>(c->qi[k].upotx);
  h.lam = make_lam(c);
  h.h2o2m = c->qi[k].h2o2mimp;
  return h;
}

// x+y Laplacians the fused step kernels start from (psi, and psix with a quantum impurity)
int prepare_lap(hedft_ctx* c) {
  if (c->lap_fresh || c->lap25_tz > 0) return HEDFT_OK;
  HEDFT_TRY(lap_xy(c, c->psi, 1));
  for (int k = 0; k < c->quantum; ++k) HEDFT_TRY(lap_xy(c, c->qi[k].psix, 1, c->qi[k].lapx));
  return HEDFT_OK;
}

// one fused H psi + integrator pass over a complex field: the 2.5-D kernel
// (x, y and z second derivatives in one pass) or the z pass after the
// separate x+y Laplacian pass.  The 2.5-D kernel reads psi at the halos of
// neighbouring tiles while the pass runs, so it writes the new psi to the
// spare buffer (the one the separate pass would hold the Laplacian in), which
// then becomes psi; a PC iteration swaps twice and an RK step four times.
template <class Out>
static void launch_step(hedft_ctx* c, void*& psi, void*& spare, Out o) {
  if (c->lap25_tz > 0) {
    o.psi = cz(spare);
    launch_lap25(c->stream, c->d, c->coef2_dev, cz(psi), o, c->lap25_tz);
    std::swap(psi, spare);
  } else {
    launch_pencil<double2, 2>(c->stream, c->d, c->coef[1][2], InPlain<double2>{cz(psi)}, o);
  }
}
static size_t step_nblocks(hedft_ctx* c) {
  return c->lap25_tz > 0 ? lap25_geom(c->d, c->lap25_tz).nblocks : pencil_geom<double2>(c->d, 2).nblocks;
}

// kernel-duration probe around one fused stepper launch (slot 0 pred, 1 corr)
static void kt_mark(hedft_ctx* c, int slot, int end) {
  if (!c->kt_on || c->kt_n[slot] >= c->kt_cap) return;
  hipEventRecord(c->kt_ev[slot][2 * c->kt_n[slot] + end], c->stream);
  if (end) c->kt_n[slot]++;
}

// RK-Gill stage of quantum impurity k (Dynamics_quantum_impurities/steprk.f90:84-106)
template <typename R>
static void steprk_x(hedft_ctx* c, double dt, int jrun, int k) {
  hedft_ctx::QImp& x = c->qi[k];
  OutRK<R, HpsiX<R>> o;
  o.partials = nullptr;
  o.h = make_hpsix<R>(c, k);
  o.psi = cz(x.psix);
  o.q = cz(x.qx);
  o.po0 = cz(x.psixold[0]);
  o.po1 = cz(x.psixold[1]);
  o.po2 = cz(x.psixold[2]);
  o.ho0 = cz(x.hpsixold[0]);
  o.ho1 = cz(x.hpsixold[1]);
  o.den = nullptr;
  o.a = RK_A[jrun];
  o.b = RK_B[jrun];
  o.cc = RK_C[jrun];
  o.dt = dt;
  o.first = (jrun == 0);
  launch_step(c, x.psix, x.lapx, o);
}

// STEPRK (Dynamics_TDDFT/steprk.f90:1-144; with quantum impurities
// Dynamics_quantum_impurities/steprk.f90:53-124: every psix takes the same stage)
template <typename R>
static int steprk_t(hedft_ctx* c, double dt) {
  for (int jrun = 0; jrun < 4; ++jrun) {
    HEDFT_TRY(prepare_lap(c));
    OutRK<R> o;
    o.partials = nullptr;
    o.h = make_hpsic<R>(c);
    o.psi = cz(c->psi);
    o.q = cz(c->q);
    o.po0 = cz(c->psiold[0]);
    o.po1 = cz(c->psiold[1]);
    o.po2 = cz(c->psiold[2]);
    o.ho0 = cz(c->hpsiold[0]);
    o.ho1 = cz(c->hpsiold[1]);
    o.den = fld<R>(c->den);
    o.a = RK_A[jrun];
    o.b = RK_B[jrun];
    o.cc = RK_C[jrun];
    o.dt = dt;
    o.first = (jrun == 0);
    launch_step(c, c->psi, c->lap, o);
    for (int k = 0; k < c->quantum; ++k) steprk_x<R>(c, dt, jrun, k);
    c->lap_fresh = 0;
    if (c->nimp > 0)
      hipLaunchKernelGGL(k_imp_rk, dim3(1), dim3(64), 0, c->stream, c->imp, c->nimp, (int)(jrun == 0), RK_A[jrun],
                         RK_B[jrun], RK_C[jrun], dt);
    if (jrun <= 2) HEDFT_TRY(poten_tddft(c, 0, 1));
  }
  for (int k = 0; k < 3; ++k) c->ioldp[k] = c->ioldr[k] = c->ioldv[k] = k;
  for (int k = 0; k < 2; ++k) c->ioldh[k] = c->iolda[k] = k;
  c->rk_done++;
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

// STEPPC (Dynamics_TDDFT/steppc.f90:1-180); errHe sum into scal[S_ERRPC].
// Quantum impurities take the same predictor and corrector
// (Dynamics_quantum_impurities/steppc.f90:64-95,115-139) with the same slot
// indices (ioldpx, ioldhx rotate with ioldp, ioldh for every impurity);
// sum|pcx_k| into scal[S_ERRPCX + k].
template <typename R>
static int steppc_t(hedft_ctx* c, double dt) {
  HEDFT_TRY(prepare_lap(c));
  {
    OutPred<R> o;
    o.partials = nullptr;
    o.h = make_hpsic<R>(c);
    o.psi = cz(c->psi);
    o.pc = cz(c->pc);
    o.po3 = cz(c->psiold[c->ioldp[2]]);
    o.ho1 = cz(c->hpsiold[c->ioldh[0]]);
    o.ho2 = cz(c->hpsiold[c->ioldh[1]]);
    o.den = fld<R>(c->den);
    o.dt = dt;
    kt_mark(c, 0, 0);
    launch_step(c, c->psi, c->lap, o);
    kt_mark(c, 0, 1);
    for (int k = 0; k < c->quantum; ++k) {
      hedft_ctx::QImp& x = c->qi[k];
      OutPred<R, HpsiX<R>> ox;
      ox.partials = nullptr;
      ox.h = make_hpsix<R>(c, k);
      ox.psi = cz(x.psix);
      ox.pc = cz(x.pcx);
      ox.po3 = cz(x.psixold[c->ioldp[2]]);
      ox.ho1 = cz(x.hpsixold[c->ioldh[0]]);
      ox.ho2 = cz(x.hpsixold[c->ioldh[1]]);
      ox.den = nullptr;
      ox.dt = dt;
      launch_step(c, x.psix, x.lapx, ox);
    }
    c->lap_fresh = 0;
    std::swap(c->ioldh[0], c->ioldh[1]);
  }
  if (c->nimp > 0) {  // runs at the start of the mean field's side branch
    c->side_pre |= 1;
    c->side_dt = dt;
    const int a[9] = {c->ioldr[2], c->ioldv[0], c->ioldv[1], c->ioldv[2], c->iolda[0], c->iolda[1], 0, 0, 0};
    std::memcpy(c->side_args, a, sizeof a);
    std::swap(c->iolda[0], c->iolda[1]);
  }
  HEDFT_TRY(poten_tddft(c, 0, 1));
  HEDFT_TRY(prepare_lap(c));
  {
    OutCorr<R> o;
    o.partials = gen_partials(c);
    o.h = make_hpsic<R>(c);
    o.psi = cz(c->psi);
    o.pc = cz(c->pc);
    o.po3 = cz(c->psiold[c->ioldp[2]]);
    o.po2 = cz(c->psiold[c->ioldp[1]]);
    o.ho1 = cz(c->hpsiold[c->ioldh[0]]);
    o.ho2 = cz(c->hpsiold[c->ioldh[1]]);
    o.den = fld<R>(c->den);
    o.dt = dt;
    kt_mark(c, 1, 0);
    launch_step(c, c->psi, c->lap, o);
    kt_mark(c, 1, 1);
    for (int k = 0; k < c->quantum; ++k) {
      hedft_ctx::QImp& x = c->qi[k];
      OutCorr<R, HpsiX<R>> ox;
      ox.partials = x_partials(c, k);
      ox.h = make_hpsix<R>(c, k);
      ox.psi = cz(x.psix);
      ox.pc = cz(x.pcx);
      ox.po3 = cz(x.psixold[c->ioldp[2]]);
      ox.po2 = cz(x.psixold[c->ioldp[1]]);
      ox.ho1 = cz(x.hpsixold[c->ioldh[0]]);
      ox.ho2 = cz(x.hpsixold[c->ioldh[1]]);
      ox.den = nullptr;
      ox.dt = dt;
      launch_step(c, x.psix, x.lapx, ox);
      c->side_pre |= 8;  // errpsix reductions, with errHe's
    }
    c->lap_fresh = 0;
    c->side_pre |= 4;  // errHe reduction, at the start of the next side branch
    c->side_nblocks = step_nblocks(c);
    const int t = c->ioldp[2];
    c->ioldp[2] = c->ioldp[1];
    c->ioldp[1] = c->ioldp[0];
    c->ioldp[0] = t;
  }
  if (c->nimp > 0) {  // runs at the start of the after-step mean field's side branch
    c->side_pre |= 2;
    c->side_dt = dt;
    const int a[9] = {c->ioldr[1], c->ioldr[2], c->ioldv[0], c->ioldv[1], c->ioldv[2], c->iolda[0], c->iolda[1],
                      0, 0};
    std::memcpy(c->side_args, a, sizeof a);
    int t = c->ioldr[2];
    c->ioldr[2] = c->ioldr[1];
    c->ioldr[1] = c->ioldr[0];
    c->ioldr[0] = t;
    t = c->ioldv[2];
    c->ioldv[2] = c->ioldv[1];
    c->ioldv[1] = c->ioldv[0];
    c->ioldv[0] = t;
  }
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

// driver part after each step (DFT4He3d.f90:686-696): potenimp, poten, forceimp,
// aimp = F/m -- the force reduction and aimp are fused into poten's gradient pass
// driver's poten + forceimp after a step; also prepares the x+y Laplacian of the
// new psi for the next step
int tddft_after_step(hedft_ctx* c) { return poten_tddft(c, 1, 1); }

int steprk_impl(hedft_ctx* c, double dt) { return c->mixed ? steprk_t<float>(c, dt) : steprk_t<double>(c, dt); }
int steppc_impl(hedft_ctx* c, double dt) { return c->mixed ? steppc_t<float>(c, dt) : steppc_t<double>(c, dt); }

}  // namespace hedft

extern "C" {

int hedft_update_aimp(hedft_ctx* c) {
  if (!c) return HEDFT_EINVAL;
  if (c->nimp > 0) hipLaunchKernelGGL(k_imp_aimp, dim3(1), dim3(1), 0, c->stream, c->imp, c->nimp);
  return done(c);
}

}  // extern "C"
