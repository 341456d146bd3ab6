// hedft.hip — mean field, integrators, impurities, energies and the C-ABI.
//
// Every entry point replaces one reference subroutine (file:line cited at the
// definition) and keeps the reference's module-state semantics: fields live in
// the context on the device; uimp and the TDDFT timec are evaluated on the fly
// from the impurity tables / absorbing-layer parameters instead of being
// stored.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>
#include <utility>

#include "stencil_kernel.h"

using namespace hedft;

namespace hedft {

// ---- device scalar slots ------------------------------------------------------
enum {
  S_MU = 0,      // sum(psi*hpsi)*dxyz/n4 of the current stage
  S_SUMSQ = 1,   // sum(psi^2)
  S_ERRPC = 2,   // sum|pc|
  S_HMEAN = 3,
  S_MUOLD = 4,
  S_ERRIMP = 5,
  S_ERRVIMP = 6,
  S_ERRPCX = 7,  // sum|pcx| of a quantum impurity
  S_ENERGY = 8,  // 8 partial energy sums
  S_ENERGYQ = 16,  // 2 quantum-impurity energy sums
  S_OBS = 16,      // 13 observable sums (used alone)
  NSCAL = 32
};

static const double RK_A[4] = {0.5, 1.0 - 0.70710678118654752440, 1.0 + 0.70710678118654752440, 1.0 / 6.0};
static const double RK_B[4] = {2.0, 1.0, 1.0, 2.0};
static const double RK_C[4] = {0.5, 1.0 - 0.70710678118654752440, 1.0 + 0.70710678118654752440, 0.5};
static const double C112 = 112.0 / 121.0, C9 = 9.0 / 121.0, C4O3 = 4.0 / 3.0;

static size_t partials_region(hedft_ctx* c) { return c->partials_cap / 8; }
static double* gen_partials(hedft_ctx* c) { return c->partials; }
static double* x_partials(hedft_ctx* c) { return c->partials + 2 * partials_region(c); }  // psix corrector

// ---- reductions -------------------------------------------------------------------
template <int NR>
__global__ void __launch_bounds__(256) k_finalize(const double* __restrict__ part, int nblocks, double* out,
                                                  double scale) {
  __shared__ double sh[NR][256];
  for (int r = 0; r < NR; ++r) {
    double s = 0.0, comp = 0.0;  // Neumaier compensated sum over this thread's blocks
    for (int b = threadIdx.x; b < nblocks; b += 256) {
      const double v = part[(size_t)b * NR + r];
      const double t = s + v;
      comp += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    sh[r][threadIdx.x] = s + comp;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < NR) out[threadIdx.x] = scale * sh[threadIdx.x][0];
}

template <int NR>
static void finalize(hedft_ctx* c, const double* part, size_t nblocks, int slot, double scale) {
  hipLaunchKernelGGL((k_finalize<NR>), dim3(1), dim3(256), 0, c->stream, part, (int)nblocks,
                     c->scal + slot, scale);
}

// ---- absorbing layer: timec = Lambdah*(1+tanh(..x)+1+tanh(..y)+1+tanh(..z)) + i
// (DFT4He3d.f90:592-601); per-axis tanh tables lam[ax][i]
struct Lam {
  const double *tx, *ty, *tz;
  double lambdah;
  // The reference assigns cmplx(Lambda,1.d0) without a KIND argument, i.e. a
  // default-kind (REAL(4)) complex, to timec: Lambda is rounded to single
  // precision before it enters the complex(8) array.
  __device__ __forceinline__ double at(int ix, int iy, int iz) const { return of(tx[ix], ty[iy], tz[iz]); }
  // from the three table values
  __device__ __forceinline__ double of(double ax, double ay, double az) const {
    return (double)(float)(lambdah * (1.0 + ax + 1.0 + ay + 1.0 + az));
  }
};

// raw operands of H psi at a point: the potential and the absorber's table
// values, loaded ahead of the arithmetic by the 2.5-D stepper kernel
struct HOps {
  double p, ax, ay, az;
  double t0, t1, rm;  // one classical impurity: its two table values and weight
};

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

// ---- energy density (energy.f90): |d psi|^2 accumulated over axes -----------------
template <typename V>
__device__ __forceinline__ double vsq(V v);
template <>
__device__ __forceinline__ double vsq<double>(double v) { return v * v; }
template <>
__device__ __forceinline__ double vsq<double2>(double2 v) { return v.x * v.x + v.y * v.y; }

template <typename V>
struct OutSq {
  static constexpr int NR = 0;
  double* partials;
  double* out;
  const double* acc;
  __device__ __forceinline__ void store(size_t i, int, int, int, V v, double*) const {
    out[i] = (acc ? acc[i] : 0.0) + vsq<V>(v);
  }
};

template <typename V>
struct OutSqXY {
  static constexpr int NR = 0;
  double* partials;
  double* out;
  __device__ __forceinline__ void store(size_t i, int, int, int, V dx, V dy, double*) const {
    out[i] = vsq<V>(dx) + vsq<V>(dy);
  }
};

template <typename V, typename R>
struct OutEnergy {
  static constexpr int NR = 8;
  double* partials;
  const double* acc;
  const R *den, *dencg, *delj4, *denalf, *dxden, *dyden, *dzden, *intx, *inty, *intz, *uext;
  ImpDev im;
  double a1, a2, den0s, beta, den_m;
  int otc;
  __device__ __forceinline__ void store(size_t i, int ix, int iy, int iz, V v, double* red) const {
    const double d = den[i], cg = dencg[i];
    red[0] += acc[i] + vsq<V>(v);
    red[1] += (double)delj4[i] * d;
    red[2] += d * cg * cg * (a1 + a2 * cg);
    if (otc)
      red[3] += (1.0 - (double)denalf[i] / den0s) * ((double)dxden[i] * intx[i] + (double)dyden[i] * inty[i] +
                                                     (double)dzden[i] * intz[i]);
    red[4] += d * (1.0 + tanh(beta * (d - den_m)));
    red[5] += (double)uext[i] * d;
    if (im.nimp > 0) red[6] += imp_u(im, ix, iy, iz) * d;
    red[7] += d;
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

// ---- statics impurity wave function (He_statics/evolox.f90) ----------------------
__global__ void k_sq(long long n, const double* a, double* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = a[i] * a[i];
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

// ---- field evaluation kernels for downloads -------------------------------------------
__global__ void k_eval_uimp(Dims d, ImpDev im, double* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < d.n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % d.nx), iy = (int)((i / d.nx) % d.ny), iz = (int)(i / ((long long)d.nx * d.ny));
    out[i] = (im.nimp > 0) ? imp_u(im, ix, iy, iz) : 0.0;
  }
}
__global__ void k_eval_timec(Dims d, Lam lam, double2* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < d.n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % d.nx), iy = (int)((i / d.nx) % d.ny), iz = (int)(i / ((long long)d.nx * d.ny));
    out[i] = make_double2(lam.at(ix, iy, iz), 1.0);
  }
}
template <typename R>
__global__ void k_eval_falfs(long long n, const R* denalf, double den0s, double* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = 1.0 - denalf[i] / den0s;
}
// TDDFT start: psi = sqrt(den) (real), den = |psi|^2 (DFT4He3d.f90:541-556)
template <typename R>
__global__ void k_tddft_psi0(long long n, R* den, double2* psi) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double s = sqrt((double)den[i]);
    psi[i] = make_double2(s, 0.0);
    den[i] = (R)(s * s);
  }
}
// precision conversion for uploads / downloads of FP32-stored fields
template <typename A, typename B>
__global__ void k_convert(long long n, const A* in, B* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (B)in[i];
}
// denx = psix*conjg(psix) (Dynamics_quantum_impurities/steprk.f90:105)
__global__ void k_abs2(long long n, const double2* z, double* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 v = z[i];
    out[i] = v.x * v.x + v.y * v.y;
  }
}
// sums of |psix|^2 and |psix|^2 (x,y,z) (DFT4He3d.f90:682-700 of the quantum driver)
__global__ void __launch_bounds__(256) k_centroid(Dims d, const double2* z, double x0, double y0, double z0,
                                                  double hx, double hy, double hz, double* partials) {
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < d.n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % d.nx);
    const long long r = i / d.nx;
    const int iy = (int)(r % d.ny), iz = (int)(r / d.ny);
    const double2 v = z[i];
    const double w = v.x * v.x + v.y * v.y;
    red[0] += w;
    red[1] += w * (x0 + hx * ix);
    red[2] += w * (y0 + hy * iy);
    red[3] += w * (z0 + hz * iz);
  }
  block_reduce_store<4>(red, partials);
}
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

// ---- kernel tables on the host: fforma / initcg / kalfs ---------------------------------
// The LJ form factor is a difference of two nearly equal terms at large |p|
// (fforma.f90:96-100), so it is evaluated in long double (x87, 64-bit mantissa)
// and rounded once.
// Sine integral Si(x): power series (Abramowitz & Stegun 5.2.14) for |x| <= 4, else
// Si = pi/2 + Im E1(i x) with E1 from its continued fraction (A&S 5.1.22, modified
// Lentz).  comp returns pi/2 - Si(x) for x > 0 without cancellation.
typedef long double ldbl;
static void sine_integral(ldbl x, ldbl* si, ldbl* comp) {
  const ldbl halfpi = 1.570796326794896619231321691639751442L;
  const ldbl ax = fabsl(x);
  if (ax <= 4.0L) {
    const ldbl x2 = ax * ax;
    ldbl term = ax, sum = ax;
    for (int k = 1; k < 80; ++k) {
      term *= -x2 / ((2.0L * k) * (2.0L * k + 1.0L));
      const ldbl t = term / (2.0L * k + 1.0L);
      sum += t;
      if (fabsl(t) < 1e-22L * fabsl(sum)) break;
    }
    const ldbl s = (x < 0) ? -sum : sum;
    *si = s;
    *comp = halfpi - s;
    return;
  }
  // continued fraction of E1(z), z = i ax, on (re, im) pairs
  ldbl br = 1.0L, bi = ax, cr = 1e300L, ci = 0.0L;
  ldbl den = br * br + bi * bi;
  ldbl dr = br / den, di = -bi / den, hr = dr, hi = di;
  for (int i = 1; i < 400; ++i) {
    const ldbl a = -(ldbl)i * i;
    br += 2.0L;
    const ldbl tr = a * dr + br, ti = a * di + bi;
    den = tr * tr + ti * ti;
    dr = tr / den;
    di = -ti / den;
    den = cr * cr + ci * ci;
    cr = br + a * cr / den;
    ci = bi - a * ci / den;
    const ldbl delr = cr * dr - ci * di, deli = cr * di + ci * dr;
    const ldbl nhr = hr * delr - hi * deli, nhi = hr * deli + hi * delr;
    hr = nhr;
    hi = nhi;
    if (fabsl(delr - 1.0L) + fabsl(deli) < 1e-21L) break;
  }
  const ldbl fi = hi * cosl(ax) - hr * sinl(ax);  // Im(h e^{-i ax}) = Si - pi/2
  const ldbl s = halfpi + fi;
  *si = (x < 0) ? -s : s;
  *comp = (x < 0) ? halfpi + s : -fi;
}

static ldbl sinx11(ldbl a, ldbl x) {  // fforma.f90:73-102: int_x^inf sin(a t)/t^11 dt
  const ldbl y = a * x, y2 = y * y;
  const ldbl b[5] = {40320.0L, -720.0L, 24.0L, -2.0L, 1.0L};
  const ldbl cc[5] = {362880.0L, -5040.0L, 120.0L, -6.0L, 1.0L};
  const ldbl d[5] = {1.0L, y2, y2 * y2, y2 * y2 * y2, y2 * y2 * y2 * y2};
  ldbl bd = 0.0L, cd = 0.0L;
  for (int i = 0; i < 5; ++i) {
    bd += b[i] * d[i];
    cd += cc[i] * d[i];
  }
  const ldbl res1 = (a * bd * cosl(y) + cd * sinl(y) / x) / (3628800.0L * powl(x, 9));
  ldbl si, comp;
  sine_integral(y, &si, &comp);
  const ldbl sinint = (a > 0.0L) ? comp : 1.570796326794896619231321691639751442L + si;
  return res1 - powl(a, 10) * sinint / 3628800.0L;
}

static ldbl sinx5(ldbl a, ldbl x) {  // fforma.f90:108-137: int_x^inf sin(a t)/t^5 dt
  const ldbl y = a * x, y2 = y * y;
  const ldbl res1 = -(a * (-2.0L + y2) * cosl(y) + (-6.0L + y2) * sinl(y) / x) / (24.0L * x * x * x);
  ldbl si, comp;
  sine_integral(y, &si, &comp);
  const ldbl sinint = (a > 0.0L) ? comp : 1.570796326794896619231321691639751442L + si;
  return res1 + powl(a, 4) * sinint / 24.0L;
}

static double fvlj_radial(double p, double eps, double sigma, double hc) {  // fforma.f90:36-60, 'OT'
  const ldbl s6 = powl((ldbl)sigma, 6);
  if (std::fabs(p) < 1e-10) {  // bforce (fforma.f90:256-280)
    const ldbl pi4 = 4.0L * 3.141592653589793238462643383279502884L, eps4 = 4.0L * eps;
    return (double)(pi4 * eps4 * s6 * s6 / (9.0L * powl((ldbl)hc, 9)) - pi4 * eps4 * s6 / (3.0L * powl((ldbl)hc, 3)));
  }
  const ldbl a = 2.0L * 3.141592653589793238462643383279502884L * (ldbl)p;
  return (double)(8.0L * (ldbl)eps * s6 * (s6 * sinx11(a, (ldbl)hc) - sinx5(a, (ldbl)hc)) / (ldbl)p);
}

static void kernel_tables_compute(const hedft_ctx* c, std::vector<double>& fv, std::vector<double>& wc,
                                  std::vector<double>& ka) {
  const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz, nxh = c->d.nxh;
  const double hpx = 1.0 / (nx * c->hx), hpy = 1.0 / (ny * c->hy), hpz = 1.0 / (nz * c->hz);
  const double pmx = 1.0 / (2.0 * c->hx), pmy = 1.0 / (2.0 * c->hy), pmz = 1.0 / (2.0 * c->hz);
  auto pgrid = [](int n, double hp, double pmax, std::vector<double>& p) {  // BCN4HeDFT.f90:337-356
    p.resize(n);
    for (int i = 1; i <= n / 2 + 1; ++i) p[i - 1] = hp * (i - 1);
    for (int i = n / 2 + 2; i <= n; ++i) p[i - 1] = -pmax + hp * (i - (n / 2 + 1));
  };
  std::vector<double> px, py, pz;
  pgrid(nx, hpx, pmx, px);
  pgrid(ny, hpy, pmy, py);
  pgrid(nz, hpz, pmz, pz);
  fv.resize(c->d.nh);
  wc.resize(c->d.nh);
  ka.resize(c->d.nh);
  const double h4 = c->p.h4, twopi = 2.0 * M_PI;
  auto planes = [&](int z0, int z1) {
    for (int iz = z0; iz < z1; ++iz)
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nxh; ++ix) {
          const size_t k = (size_t)ix + (size_t)nxh * (iy + (size_t)ny * iz);
          const double pm = std::sqrt(px[ix] * px[ix] + py[iy] * py[iy] + pz[iz] * pz[iz]);
          fv[k] = fvlj_radial(pm, c->p.eps4, c->p.sigma4, h4);
          double w = 1.0;  // initcg.f90:32-49
          if (pm > 0.0) {
            const double arg = twopi * pm * h4;
            const double aux = std::sin(arg) - arg * std::cos(arg);
            w = (std::fabs(aux) < 1e-10) ? 1.0 : 3.0 * aux / (arg * arg * arg);
          }
          wc[k] = w;
          const double e = M_PI * c->p.l * pm;  // BCN4HeDFT.f90:849
          ka[k] = std::exp(-e * e);
        }
  };
  // z-planes split over host threads (each value depends only on its own p)
  const int nt = std::max(1, std::min({(int)std::thread::hardware_concurrency(), 16, nz}));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t) pool.emplace_back(planes, nz * t / nt, nz * (t + 1) / nt);
  for (auto& th : pool) th.join();
}

// The tables depend only on the grid and the functional constants; contexts
// with the same ones (an ensemble, a serving loop) share one host computation.
static void kernel_tables(const hedft_ctx* c, std::vector<double>& fv, std::vector<double>& wc,
                          std::vector<double>& ka) {
  struct Entry {
    double key[10];
    std::vector<double> fv, wc, ka;
  };
  static std::mutex mu;
  static std::vector<Entry> cache;
  const double key[10] = {(double)c->d.nx, (double)c->d.ny, (double)c->d.nz, c->hx, c->hy, c->hz,
                          c->p.h4, c->p.eps4, c->p.sigma4, c->p.l};
  {
    std::lock_guard<std::mutex> lk(mu);
    for (const Entry& e : cache)
      if (std::memcmp(e.key, key, sizeof key) == 0) {
        fv = e.fv;
        wc = e.wc;
        ka = e.ka;
        return;
      }
  }
  kernel_tables_compute(c, fv, wc, ka);
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 8) cache.erase(cache.begin());
  Entry e;
  std::memcpy(e.key, key, sizeof key);
  e.fv = fv;
  e.wc = wc;
  e.ka = ka;
  cache.push_back(std::move(e));
}

// ---- mean field pipelines --------------------------------------------------------------------
// ---- two-branch sections: work launched between fork() and join() inside a
// SideBranch scope runs on the side stream, concurrently with what the main
// stream does in the meantime; join() makes the main stream wait for it.  The
// branches must touch disjoint buffers.  Under stream capture the same calls
// record a graph with two parallel branches.
static int fork(hedft_ctx* c) {
  if (c->no_side) return HEDFT_OK;
  HEDFT_CHECK(hipEventRecord(c->ev_fork, c->main_stream));
  HEDFT_CHECK(hipStreamWaitEvent(c->side_stream, c->ev_fork, 0));
  return HEDFT_OK;
}

// A fork in two halves: fork_point() marks the branch point on the main
// stream, the main branch's first kernel is launched, then fork_side() lets the
// side stream start.  Both branches still begin at the branch point, but the
// main kernel (a whole-plane FFT that needs a full CU's LDS) is dispatched
// before the side branch's kernels spread over the CUs.
static int fork_point(hedft_ctx* c) {
  if (c->no_side) return HEDFT_OK;
  HEDFT_CHECK(hipEventRecord(c->ev_fork, c->main_stream));
  return HEDFT_OK;
}
static int fork_side(hedft_ctx* c) {
  if (c->no_side) return HEDFT_OK;
  HEDFT_CHECK(hipStreamWaitEvent(c->side_stream, c->ev_fork, 0));
  return HEDFT_OK;
}

static int join(hedft_ctx* c) {
  if (c->no_side) return HEDFT_OK;
  HEDFT_CHECK(hipEventRecord(c->ev_join, c->side_stream));
  HEDFT_CHECK(hipStreamWaitEvent(c->main_stream, c->ev_join, 0));
  return HEDFT_OK;
}

struct SideBranch {
  hedft_ctx* c;
  explicit SideBranch(hedft_ctx* cc) : c(cc) { c->stream = c->no_side ? c->main_stream : c->side_stream; }
  ~SideBranch() { c->stream = c->main_stream; }
};

// field pointers as stored: the double* members hold float storage on the mixed path
template <typename R>
static inline R* fld(double* p) {
  return reinterpret_cast<R*>(p);
}
template <typename R>
static inline cplx_t<R>* spc(double2* p) {
  return reinterpret_cast<cplx_t<R>*>(p);
}

template <typename R>
static XOutT<R> xo_store(double* out) {
  XOutT<R> o{};
  o.op = XOUT_STORE;
  o.out = fld<R>(out);
  return o;
}

template <typename R>
static XOutT<R> xo_pot4(hedft_ctx* c) {
  XOutT<R> o{};
  o.op = XOUT_POT4;
  o.out = fld<R>(c->pot4);
  o.delj4 = fld<R>(c->delj4);
  o.dencg = fld<R>(c->dencg);
  o.den = fld<R>(c->den);
  o.ext = fld<R>(pot4_ext(c));
  o.a0 = c->p.cp4 / 2.0;
  o.a1 = c->p.cpp4 / 3.0;
  o.csolid = c->p.csolid;
  o.beta = c->p.beta_solid;
  o.den_m = c->p.den_m;
  o.lsolid = c->p.lsolid;
  o.lext = pot4_ext(c) != nullptr;
  return o;
}

// Stage B of the mean field, one batched convolution of up to four arrays:
//   corr  = F^-1[wcgk F(den dencg (cp4+cpp4 dencg))]          (poten.f90:166-178)
//   int_a = F^-1[kalfs F((1-denalf/den0s) d_a den)], a = x,y,z  (term_alfa.f90:26-48)
// For the OT core only corr is needed and its C2R epilogue assembles pot4.
template <typename R>
static int stage_b(hedft_ctx* c) {
  XInT<R> in[4];
  in[0] = XInT<R>{};
  in[0].op = XIN_CORR;
  in[0].a[0] = fld<R>(c->den);
  in[0].a[1] = fld<R>(c->dencg);
  in[0].c0 = c->p.cp4;
  in[0].c1 = c->p.cpp4;
  const R* grads[3] = {fld<R>(c->dxden), fld<R>(c->dyden), fld<R>(c->dzden)};
  for (int a = 0; a < 3; ++a) {
    in[1 + a] = XInT<R>{};
    in[1 + a].op = XIN_FGRAD;
    in[1 + a].a[0] = fld<R>(c->denalf);
    in[1 + a].a[1] = grads[a];
    in[1 + a].c0 = c->p.den0s;
  }
  const int nb = c->p.core_otc ? 4 : 1;
  cplx_t<R>* sp[4] = {spc<R>(c->spec[0]), spc<R>(c->spec[1]), spc<R>(c->spec[2]), spc<R>(c->spec[3])};
  HEDFT_TRY(fwd_xy<R>(c, nb, in, sp));
  const R* kk[4] = {fld<R>(c->wcgk), fld<R>(c->kalfs), fld<R>(c->kalfs), fld<R>(c->kalfs)};
  HEDFT_TRY(zconv<R>(c, nb, sp, 1, kk, sp, 1, nullptr));
  XOutT<R> xo[4] = {c->p.core_otc ? xo_store<R>(c->sto1) : xo_pot4<R>(c), xo_store<R>(c->intx),
                    xo_store<R>(c->inty), xo_store<R>(c->intz)};
  return inv_yx<R>(c, nb, sp, xo);
}

// Stage C: ualphas' convolution F^-1[kalfs F(grad(den).int)] (term_alfa.f90:61-73),
// then the divergence term and the pot4 assembly (term_alfa.f90:80-98, poten.f90:194-211)
// the single-stream mean field (tuning field `fused`): with the 2.5-D steppers
// on grids the 2.5-D tiles divide
static bool fused_mf(const hedft_ctx* c) { return c->fused && c->lap25_tz > 0 && div25_ok(c); }

template <typename R>
static int stage_c(hedft_ctx* c) {
  const bool fused = fused_mf(c);
  // side branch: x+y divergence (denalf, int_x, int_y -> work1); main branch:
  // the convolution (grad den, int -> spec[0] -> ualph_conv).  Fused: the
  // divergence is one 2.5-D pass after the convolution, on the main stream.
  if (!fused) HEDFT_TRY(fork_point(c));
  XInT<R> dot{};
  dot.op = XIN_DOT3;
  dot.a[0] = fld<R>(c->dxden);
  dot.a[1] = fld<R>(c->dyden);
  dot.a[2] = fld<R>(c->dzden);
  dot.a[3] = fld<R>(c->intx);
  dot.a[4] = fld<R>(c->inty);
  dot.a[5] = fld<R>(c->intz);
  cplx_t<R>* s0 = spc<R>(c->spec[0]);
  HEDFT_TRY(fwd_xy<R>(c, 1, &dot, &s0));
  if (!fused) {
    HEDFT_TRY(fork_side(c));
    SideBranch sb(c);
    HEDFT_TRY(alpha_div_xy(c));
  }
  const R* ka = fld<R>(c->kalfs);
  HEDFT_TRY(zconv<R>(c, 1, &s0, 1, &ka, &s0, 1, nullptr));
  // ualphas' convolution, with the density convolutions the mean field left
  // for this inverse (delj4 [, upotx])
  cplx_t<R>* su[3] = {s0, nullptr, nullptr};
  XOutT<R> xu[3] = {xo_store<R>(c->ualph_conv), {}, {}};
  int nu = 1;
  for (int k = 0; k < c->ndefer; ++k) {
    su[nu] = spc<R>(c->defer_spec[k]);
    xu[nu++] = xo_store<R>(c->defer_out[k]);
  }
  c->ndefer = 0;
  HEDFT_TRY(inv_yx<R>(c, nu, su, xu));
  if (fused) return alpha_div25_pot4(c);
  HEDFT_TRY(join(c));
  return alpha_div_z_pot4(c);
}

// term_alfa() on its own (B restricted to int_a, then C)
template <typename R>
static int alpha_term(hedft_ctx* c) {
  XInT<R> fg[3];
  const R* grads[3] = {fld<R>(c->dxden), fld<R>(c->dyden), fld<R>(c->dzden)};
  for (int a = 0; a < 3; ++a) {
    fg[a] = XInT<R>{};
    fg[a].op = XIN_FGRAD;
    fg[a].a[0] = fld<R>(c->denalf);
    fg[a].a[1] = grads[a];
    fg[a].c0 = c->p.den0s;
  }
  cplx_t<R>* s3[3] = {spc<R>(c->spec[1]), spc<R>(c->spec[2]), spc<R>(c->spec[3])};
  HEDFT_TRY(fwd_xy<R>(c, 3, fg, s3));
  const R* ka[3] = {fld<R>(c->kalfs), fld<R>(c->kalfs), fld<R>(c->kalfs)};
  HEDFT_TRY(zconv<R>(c, 3, s3, 1, ka, s3, 1, nullptr));
  XOutT<R> xo3[3] = {xo_store<R>(c->intx), xo_store<R>(c->inty), xo_store<R>(c->intz)};
  HEDFT_TRY(inv_yx<R>(c, 3, s3, xo3));
  return stage_c<R>(c);
}

// the serial kernels a step left for the side branch (impurity Newton PC,
// errHe reduction); also called directly where no mean field follows
static void flush_side_errors(hedft_ctx* c) {
  if (c->side_pre & 4)
    finalize<1>(c, gen_partials(c), c->side_nblocks, S_ERRPC, 1.0);
  if (c->side_pre & 8)
    finalize<1>(c, x_partials(c), c->side_nblocks, S_ERRPCX, 1.0);
  c->side_pre &= ~(4 | 8);
}
static void flush_side_impurity(hedft_ctx* c) {
  const int* a = c->side_args;
  if (c->side_pre & 1)
    hipLaunchKernelGGL(k_imp_pred, dim3(1), dim3(64), 0, c->stream, c->imp, c->nimp, c->side_dt, a[0], a[1], a[2],
                       a[3], a[4], a[5]);
  if (c->side_pre & 2)
    hipLaunchKernelGGL(k_imp_corr, dim3(1), dim3(64), 0, c->stream, c->imp, c->nimp, c->side_dt, a[0], a[1], a[2],
                       a[3], a[4], a[5], a[6], c->scal);
  c->side_pre &= ~(1 | 2);
}
static void flush_side_pre(hedft_ctx* c) {
  flush_side_errors(c);
  flush_side_impurity(c);
}

// potx4 = F^-1[F(|psix|^2) uimp_k] of a quantum impurity
// (Dynamics_quantum_impurities/potenimp.f90:22-27), one forward, one
// convolution and one inverse of the impurity density
template <typename R>
static int potx4_conv(hedft_ctx* c, double* out) {
  XInT<R> xa{};
  xa.op = XIN_ABS2;
  xa.z = (const double2*)c->psix;
  cplx_t<R>* sx = spc<R>(c->specq[0]);
  HEDFT_TRY(fwd_xy<R>(c, 1, &xa, &sx));
  const R* ku = fld<R>(c->uimpk);
  HEDFT_TRY(zconv<R>(c, 1, &sx, 1, &ku, &sx, 1, nullptr));
  XOutT<R> xo = xo_store<R>(out);
  return inv_yx<R>(c, 1, &sx, &xo);
}

// potenimp of Dynamics_quantum_impurities (potenimp.f90:20-30) on its own:
// upotx = F^-1[F(den) uimp_k], potx4
template <typename R>
static int potenimp_q(hedft_ctx* c) {
  XInT<R> xi{};
  xi.op = XIN_COPY;
  xi.a[0] = fld<R>(c->den);
  cplx_t<R>* s0 = spc<R>(c->spec[0]);
  HEDFT_TRY(fwd_xy<R>(c, 1, &xi, &s0));
  const R* ku = fld<R>(c->uimpk);
  cplx_t<R>* s1 = spc<R>(c->specq[1]);
  HEDFT_TRY(zconv<R>(c, 1, &s0, 1, &ku, &s1, 1, nullptr));
  XOutT<R> xo = xo_store<R>(c->upotx);
  HEDFT_TRY(inv_yx<R>(c, 1, &s1, &xo));
  return potx4_conv<R>(c, c->potx4);
}

// Dynamics_TDDFT/poten.f90:124-221 (+ forceimp partials fused into derden).
// grad(den) and the impurity force run on the side branch while the main
// branch convolves den; with_lap also computes the x+y Laplacian of psi there
// for the step that follows (psi is not changed by the mean field).
// With a quantum impurity this is potenimp + poten of
// Dynamics_quantum_impurities (potenimp.f90:20-30, poten.f90:26-138): upotx =
// F^-1[F(den) uimp_k] is one more output of the den convolution, potx4 is
// convolved on the side branch and added in the pot4 assembly.
template <typename R>
static int poten_tddft_t(hedft_ctx* c, int want_force, int with_lap) {
  // where the next step's x+y Laplacians run: beside the den convolutions
  // (early) or beside stages B and C (late); lap_late 2 puts them late only in
  // the mean field that also carries the impurity force, 3 on the main stream
  // after the den convolutions (beside the rest of the side branch), 4 on the
  // side stream after grad(den) with stage B waiting only for grad(den) (the
  // Laplacians are joined with stage C's branch, or after stage B).  With the
  // 2.5-D steppers there is no Laplacian pass; lap_late 4 still lets stage B
  // wait only for grad(den).
  // (with the fused stage C no join follows stage B, so the side branch is
  // joined whole before it)
  const bool mid = (with_lap || c->lap25_tz > 0) && c->lap_late == 4 && !c->no_side && !fused_mf(c);
  if (c->lap25_tz > 0) with_lap = 0;
  const bool late = with_lap && (c->lap_late == 1 || (c->lap_late == 2 && want_force));
  const bool on_main = with_lap && c->lap_late == 3;
  HEDFT_TRY(fork_point(c));
  XInT<R> xi{};
  xi.op = XIN_COPY;
  xi.a[0] = fld<R>(c->den);
  cplx_t<R>* s0 = spc<R>(c->spec[0]);
  HEDFT_TRY(fwd_xy<R>(c, 1, &xi, &s0));
  HEDFT_TRY(fork_side(c));
  {
    SideBranch sb(c);
    // impurity PC and errHe of the step just queued; with the mid-branch join
    // the errHe sum and the force reduction (read only by the next step's
    // impurity PC) wait until grad(den) is done
    const bool derden = !c->quantum || c->p.core_otc;
    if (!mid) flush_side_errors(c);
    flush_side_impurity(c);
    // the quantum-impurity poten calls derden only for the OTC core
    // (Dynamics_quantum_impurities/poten.f90:48): nothing else reads grad(den) there
    if (derden) HEDFT_TRY(derden_force(c, want_force, want_force, !mid));
    if (c->quantum) HEDFT_TRY(potx4_conv<R>(c, c->potx4));
    if (mid) {
      HEDFT_CHECK(hipEventRecord(c->ev_mid, c->side_stream));
      if (derden && want_force) HEDFT_TRY(derden_force_finalize(c, want_force));
      flush_side_errors(c);
    }
    if (with_lap && !late && !on_main) {
      HEDFT_TRY(lap_xy(c, c->psi, 1));
      if (c->quantum) HEDFT_TRY(lap_xy(c, c->psix, 1, c->lapx));
    }
  }
  // den convolutions: dencg, delj4 [, denalf] [, upotx].  With the OTC core,
  // delj4 (read only by the pot4 assembly at the end of stage C) and upotx
  // (read by the next psix step) are inverted with stage C's convolution, so
  // that this inverse needs one round of workgroups over the CUs instead of two
  const bool defer = c->p.core_otc && c->specd;
  const R* kk[4] = {fld<R>(c->wcgk), fld<R>(c->fvlj4), nullptr, nullptr};
  cplx_t<R>* so[4] = {spc<R>(c->spec[1]), spc<R>(defer ? c->specd : c->spec[2]), nullptr, nullptr};
  XOutT<R> xo[4] = {xo_store<R>(c->dencg), xo_store<R>(c->delj4), {}, {}};
  int nk = 2;
  if (c->p.core_otc) {
    kk[nk] = fld<R>(c->kalfs);
    so[nk] = spc<R>(c->spec[3]);
    xo[nk++] = xo_store<R>(c->denalf);
  }
  if (c->quantum) {
    kk[nk] = fld<R>(c->uimpk);
    so[nk] = spc<R>(c->specq[1]);
    xo[nk++] = xo_store<R>(c->upotx);
  }
  if (c->zconv_fused) {
    // one workgroup per column tile applies up to three kernels
    const int n1 = std::min(nk, 3);
    HEDFT_TRY(zconv<R>(c, 1, &s0, n1, kk, so, 1, nullptr));
    if (nk > n1) HEDFT_TRY(zconv<R>(c, 1, &s0, nk - n1, kk + n1, so + n1, 1, nullptr));
  } else {
    // one workgroup per (column tile, kernel): the forward z transform is
    // repeated per kernel, but nk times as many workgroups are in flight
    cplx_t<R>* si[4] = {s0, s0, s0, s0};
    HEDFT_TRY(zconv<R>(c, nk, si, 1, kk, so, 1, nullptr));
  }
  c->ndefer = 0;
  if (defer) {
    cplx_t<R>* si[4];
    XOutT<R> xi[4];
    int ni = 0;
    for (int k = 0; k < nk; ++k) {
      if (xo[k].out == fld<R>(c->delj4) || (c->quantum && xo[k].out == fld<R>(c->upotx))) {
        c->defer_spec[c->ndefer] = reinterpret_cast<double2*>(so[k]);
        c->defer_out[c->ndefer++] = reinterpret_cast<double*>(xo[k].out);
      } else {
        si[ni] = so[k];
        xi[ni++] = xo[k];
      }
    }
    HEDFT_TRY(inv_yx<R>(c, ni, si, xi));
  } else {
    HEDFT_TRY(inv_yx<R>(c, nk, so, xo));
  }
  if (on_main) {
    HEDFT_TRY(lap_xy(c, c->psi, 1));
    if (c->quantum) HEDFT_TRY(lap_xy(c, c->psix, 1, c->lapx));
  }
  if (mid) HEDFT_CHECK(hipStreamWaitEvent(c->main_stream, c->ev_mid, 0));
  else HEDFT_TRY(join(c));
  // the x+y Laplacians for the next step run beside stages B and C (they read
  // only psi and psix, which the mean field does not change); stage C's own
  // branch and its join come after them on the side stream
  if (late) {
    HEDFT_TRY(fork(c));
    {
      SideBranch sb(c);
      HEDFT_TRY(lap_xy(c, c->psi, 1));
      if (c->quantum) HEDFT_TRY(lap_xy(c, c->psix, 1, c->lapx));
    }
  }
  if (with_lap) c->lap_fresh = 1;
  HEDFT_TRY(stage_b<R>(c));
  if (c->p.core_otc) HEDFT_TRY(stage_c<R>(c));
  else if (late || mid) HEDFT_TRY(join(c));
  return HEDFT_OK;
}

static int poten_tddft(hedft_ctx* c, int want_force, int with_lap = 0) {
  return c->mixed ? poten_tddft_t<float>(c, want_force, with_lap) : poten_tddft_t<double>(c, want_force, with_lap);
}

// He_statics/poten.f90:5-221: fden, dencg, denalf and grad(den) as left by the caller
// (with an impurity wave function also fdenx: upotx = F^-1[fden vq], potx4 =
// F^-1[fdenx vq] enter the pot4 assembly, poten.f90:94-107)
static int poten_statics(hedft_ctx* c) {
  const double* kf = c->fvlj4;
  double2* s2 = c->spec[2];
  double2* fd = c->fden;
  HEDFT_TRY(zconv(c, 1, &fd, 1, &kf, &s2, 0, nullptr));
  XOut xo = xo_store<double>(c->delj4);
  HEDFT_TRY(inv_yx(c, 1, &s2, &xo));
  if (c->quantum) {
    double2* in[2] = {c->fden, c->specq[0]};
    const double* kq[2] = {c->uimpk, c->uimpk};
    double2* so[2] = {c->spec[0], c->spec[1]};
    HEDFT_TRY(zconv(c, 2, in, 1, kq, so, 0, nullptr));
    XOut xq[2] = {xo_store<double>(c->upotx), xo_store<double>(c->potx4)};
    HEDFT_TRY(inv_yx(c, 2, so, xq));
  }
  HEDFT_TRY(stage_b<double>(c));
  if (c->p.core_otc) HEDFT_TRY(stage_c<double>(c));
  return HEDFT_OK;
}

// fdenx = F(psix^2) of the statics impurity (BCN4HeDFT.f90:788,1130)
static int statics_fdenx(hedft_ctx* c) {
  hipLaunchKernelGGL(k_sq, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, (const double*)c->psix, c->work2);
  return fftfw(c, c->work2, c->specq[0]);
}

// driver refresh after a statics step (BCN4HeDFT.f90:1131-1149): derden, fden, dencg, denalf
// (and fdenx with an impurity wave function, BCN4HeDFT.f90:1128-1130)
static int statics_refresh(hedft_ctx* c) {
  if (c->quantum) HEDFT_TRY(statics_fdenx(c));
  HEDFT_TRY(fork(c));
  {
    SideBranch sb(c);
    HEDFT_TRY(derden_force(c, 0, 0));
  }
  XIn xi{};
  xi.op = XIN_COPY;
  xi.a[0] = c->den;
  double2* s0 = c->spec[0];
  HEDFT_TRY(fwd_xy(c, 1, &xi, &s0));
  double2* fd = c->fden;
  if (c->p.core_otc) {
    const double* kk[2] = {c->wcgk, c->kalfs};
    double2* so[2] = {c->spec[1], c->spec[3]};
    HEDFT_TRY(zconv(c, 1, &s0, 2, kk, so, 1, &fd));
    XOut xo[2] = {xo_store<double>(c->dencg), xo_store<double>(c->denalf)};
    HEDFT_TRY(inv_yx(c, 2, so, xo));
  } else {
    const double* kw = c->wcgk;
    double2* s1 = c->spec[1];
    HEDFT_TRY(zconv(c, 1, &s0, 1, &kw, &s1, 1, &fd));
    XOut xo = xo_store<double>(c->dencg);
    HEDFT_TRY(inv_yx(c, 1, &s1, &xo));
  }
  return join(c);
}

int poten(hedft_ctx* c, int want_force) {
  return (c->mode == HEDFT_MODE_TDDFT) ? poten_tddft(c, want_force) : poten_statics(c);
}


// ---- integrator drivers ------------------------------------------------------------------------
static Lam make_lam(hedft_ctx* c) {
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
static HpsiX<R> make_hpsix(hedft_ctx* c) {
  HpsiX<R> h;
  h.lapxy = c->lap25_tz > 0 ? nullptr : (const double2*)c->lapx;
  h.upot = fld<R>(c->upotx);
  h.lam = make_lam(c);
  h.h2o2m = c->h2o2mimp;
  return h;
}

// x+y Laplacians the fused step kernels start from (psi, and psix with a quantum impurity)
static int prepare_lap(hedft_ctx* c) {
  if (c->lap_fresh || c->lap25_tz > 0) return HEDFT_OK;
  HEDFT_TRY(lap_xy(c, c->psi, 1));
  if (c->quantum) HEDFT_TRY(lap_xy(c, c->psix, 1, c->lapx));
  return HEDFT_OK;
}

static inline double2* cz(void* p) { return (double2*)p; }
static inline double* rz(void* p) { return (double*)p; }

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

// RK-Gill stage of a quantum impurity (Dynamics_quantum_impurities/steprk.f90:84-106)
template <typename R>
static void steprk_x(hedft_ctx* c, double dt, int jrun) {
  OutRK<R, HpsiX<R>> o;
  o.partials = nullptr;
  o.h = make_hpsix<R>(c);
  o.psi = cz(c->psix);
  o.q = cz(c->qx);
  o.po0 = cz(c->psixold[0]);
  o.po1 = cz(c->psixold[1]);
  o.po2 = cz(c->psixold[2]);
  o.ho0 = cz(c->hpsixold[0]);
  o.ho1 = cz(c->hpsixold[1]);
  o.den = nullptr;
  o.a = RK_A[jrun];
  o.b = RK_B[jrun];
  o.cc = RK_C[jrun];
  o.dt = dt;
  o.first = (jrun == 0);
  launch_step(c, c->psix, c->lapx, o);
}

// STEPRK (Dynamics_TDDFT/steprk.f90:1-144; with a quantum impurity
// Dynamics_quantum_impurities/steprk.f90:53-124: psix takes the same stage)
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
    if (c->quantum) steprk_x<R>(c, dt, jrun);
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
// A quantum impurity takes the same predictor and corrector
// (Dynamics_quantum_impurities/steppc.f90:64-95,116-139) with the same slot
// indices (ioldpx, ioldhx rotate with ioldp, ioldh); sum|pcx| into scal[S_ERRPCX].
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
    if (c->quantum) {
      OutPred<R, HpsiX<R>> ox;
      ox.partials = nullptr;
      ox.h = make_hpsix<R>(c);
      ox.psi = cz(c->psix);
      ox.pc = cz(c->pcx);
      ox.po3 = cz(c->psixold[c->ioldp[2]]);
      ox.ho1 = cz(c->hpsixold[c->ioldh[0]]);
      ox.ho2 = cz(c->hpsixold[c->ioldh[1]]);
      ox.den = nullptr;
      ox.dt = dt;
      launch_step(c, c->psix, c->lapx, ox);
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
    if (c->quantum) {
      OutCorr<R, HpsiX<R>> ox;
      ox.partials = x_partials(c);
      ox.h = make_hpsix<R>(c);
      ox.psi = cz(c->psix);
      ox.pc = cz(c->pcx);
      ox.po3 = cz(c->psixold[c->ioldp[2]]);
      ox.po2 = cz(c->psixold[c->ioldp[1]]);
      ox.ho1 = cz(c->hpsixold[c->ioldh[0]]);
      ox.ho2 = cz(c->hpsixold[c->ioldh[1]]);
      ox.den = nullptr;
      ox.dt = dt;
      launch_step(c, c->psix, c->lapx, ox);
      c->side_pre |= 8;  // errpsix reduction, with errHe's
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
static int tddft_after_step(hedft_ctx* c) { return poten_tddft(c, 1, 1); }

static int steprk_impl(hedft_ctx* c, double dt) { return c->mixed ? steprk_t<float>(c, dt) : steprk_t<double>(c, dt); }
static int steppc_impl(hedft_ctx* c, double dt) { return c->mixed ? steppc_t<float>(c, dt) : steppc_t<double>(c, dt); }

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

static int read_scal(hedft_ctx* c, int first, int count, double* out) {
  HEDFT_CHECK(hipMemcpyAsync(c->host_scal, c->scal + first, count * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < count; ++i) out[i] = c->host_scal[i];
  return HEDFT_OK;
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

// ---- energy -----------------------------------------------------------------------------------
template <typename V, typename R>
static int energy_sums(hedft_ctx* c, double* s) {
  const V* psi = (const V*)c->psi;
  launch_pencil_xy<V, true>(c->stream, c->d, c->coef[0][0], c->coef[0][1], InPlain<V>{psi},
                            OutSqXY<V>{nullptr, (double*)c->workc1});
  OutEnergy<V, R> o;
  o.partials = gen_partials(c);
  o.acc = (const double*)c->workc1;
  o.den = fld<R>(c->den);
  o.dencg = fld<R>(c->dencg);
  o.delj4 = fld<R>(c->delj4);
  o.denalf = fld<R>(c->denalf);
  o.dxden = fld<R>(c->dxden);
  o.dyden = fld<R>(c->dyden);
  o.dzden = fld<R>(c->dzden);
  o.intx = fld<R>(c->intx);
  o.inty = fld<R>(c->inty);
  o.intz = fld<R>(c->intz);
  // sum(uext den), or sum(potx4 den) with a statics impurity wave function
  // (He_statics/energy.f90:91)
  o.uext = fld<R>((c->quantum && c->mode == HEDFT_MODE_STATICS) ? c->potx4 : c->uext);
  o.im = make_impdev(c);
  o.a1 = 0.5 * c->p.cp4;
  o.a2 = c->p.cpp4 / 3.0;
  o.den0s = c->p.den0s;
  o.beta = c->p.beta_solid;
  o.den_m = c->p.den_m;
  o.otc = c->p.core_otc;
  launch_pencil<V, 2>(c->stream, c->d, c->coef[0][2], InPlain<V>{psi}, o);
  hipLaunchKernelGGL(k_finalize<8>, dim3(1), dim3(256), 0, c->stream, gen_partials(c),
                     (int)pencil_geom<V>(c->d, 2).nblocks, c->scal + S_ENERGY, 1.0);
  return read_scal(c, S_ENERGY, 8, s);
}

// quantum impurity terms (Dynamics_quantum_impurities/energy.f90:102-132):
// sum |grad psix|^2 and sum F^-1[F(|psix|^2) uimp_k] den
template <typename R>
struct OutEnergyX {
  static constexpr int NR = 2;
  double* partials;
  const double* acc;
  const R *conv, *den;
  __device__ __forceinline__ void store(size_t i, int, int, int, double2 v, double* red) const {
    red[0] += acc[i] + vsq<double2>(v);
    red[1] += (double)conv[i] * (double)den[i];
  }
};

template <typename R>
static int energy_sums_x(hedft_ctx* c, double* s) {
  HEDFT_TRY(potx4_conv<R>(c, c->work1));
  const double2* px = (const double2*)c->psix;
  launch_pencil_xy<double2, true>(c->stream, c->d, c->coef[0][0], c->coef[0][1], InPlain<double2>{px},
                                  OutSqXY<double2>{nullptr, (double*)c->workc1});
  OutEnergyX<R> o;
  o.partials = gen_partials(c);
  o.acc = (const double*)c->workc1;
  o.conv = fld<R>(c->work1);
  o.den = fld<R>(c->den);
  launch_pencil<double2, 2>(c->stream, c->d, c->coef[0][2], InPlain<double2>{px}, o);
  hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(256), 0, c->stream, gen_partials(c),
                     (int)pencil_geom<double2>(c->d, 2).nblocks, c->scal + S_ENERGYQ, 1.0);
  return read_scal(c, S_ENERGYQ, 2, s);
}

// statics impurity wave function: sum |grad psix|^2 (He_statics/energy.f90:81-90)
struct OutSumSq {
  static constexpr int NR = 1;
  double* partials;
  const double* acc;
  __device__ __forceinline__ void store(size_t i, int, int, int, double v, double* red) const {
    red[0] += acc[i] + v * v;
  }
};
static int energy_sums_xs(hedft_ctx* c, double* s) {
  const double* px = rz(c->psix);
  launch_pencil_xy<double, true>(c->stream, c->d, c->coef[0][0], c->coef[0][1], InPlain<double>{px},
                                 OutSqXY<double>{nullptr, (double*)c->workc1});
  OutSumSq o;
  o.partials = gen_partials(c);
  o.acc = (const double*)c->workc1;
  launch_pencil<double, 2>(c->stream, c->d, c->coef[0][2], InPlain<double>{px}, o);
  finalize<1>(c, gen_partials(c), pencil_geom<double>(c->d, 2).nblocks, S_ENERGYQ, 1.0);
  return read_scal(c, S_ENERGYQ, 1, s);
}

// ---- droplet observables (Dynamics_TDDFT/DFT4He3d.f90:709-836, r_cm.f90) ----------------------
// x and y first derivatives of psi into two complex fields
struct OutTwoC {
  static constexpr int NR = 0;
  double2 *ox, *oy;
  double* partials;
  __device__ __forceinline__ void store(size_t i, int, int, int, double2 dx, double2 dy, double*) const {
    ox[i] = dx;
    oy[i] = dy;
  }
};

// z pass with the sums: den moments (7) and, with g = grad psi,
// Re[i conj(psi) g_a] (3) and Re[i conj(psi) (r x g)_a] about the origin (3)
template <typename R>
struct OutObs {
  static constexpr int NR = 13;
  double* partials;
  const double2 *psi, *dx, *dy;
  const R* den;
  double x0, y0, z0, hx, hy, hz;
  __device__ __forceinline__ void store(size_t i, int ix, int iy, int iz, double2 dz, double* red) const {
    const double x = x0 + hx * ix, y = y0 + hy * iy, z = z0 + hz * iz;
    const double r = (double)den[i];
    red[0] += r;
    red[1] += r * x;
    red[2] += r * y;
    red[3] += r * z;
    red[4] += r * x * x;
    red[5] += r * y * y;
    red[6] += r * z * z;
    const double2 p = psi[i], gx = dx[i], gy = dy[i];
    // Re[i conj(p) g] = p.y g.x - p.x g.y
    const double jx = p.y * gx.x - p.x * gx.y, jy = p.y * gy.x - p.x * gy.y, jz = p.y * dz.x - p.x * dz.y;
    red[7] += z * jy - y * jz;
    red[8] += x * jz - z * jx;
    red[9] += y * jx - x * jy;
    red[10] += jx;
    red[11] += jy;
    red[12] += jz;
  }
};

template <typename R>
static int observable_sums(hedft_ctx* c, double* s) {
  const double2* psi = (const double2*)c->psi;
  double2 *gx = (double2*)c->workc1, *gy = (double2*)c->lap;
  launch_pencil_xy<double2, true>(c->stream, c->d, c->coef[0][0], c->coef[0][1], InPlain<double2>{psi},
                                  OutTwoC{gx, gy, nullptr});
  c->lap_fresh = 0;  // the Laplacian buffer held gy
  OutObs<R> o;
  o.partials = gen_partials(c);
  o.psi = psi;
  o.dx = gx;
  o.dy = gy;
  o.den = fld<R>(c->den);
  o.x0 = -c->g.xmax;
  o.y0 = -c->g.ymax;
  o.z0 = -c->g.zmax;
  o.hx = c->hx;
  o.hy = c->hy;
  o.hz = c->hz;
  launch_pencil<double2, 2>(c->stream, c->d, c->coef[0][2], InPlain<double2>{psi}, o);
  hipLaunchKernelGGL(k_finalize<13>, dim3(1), dim3(256), 0, c->stream, gen_partials(c),
                     (int)pencil_geom<double2>(c->d, 2).nblocks, c->scal + S_OBS, 1.0);
  return read_scal(c, S_OBS, 13, s);
}

}  // namespace hedft

// ---------------------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------------------
extern "C" {

const char* hedft_strerror(int rc) {
  switch (rc) {
    case HEDFT_OK: return "ok";
    case HEDFT_EINVAL: return "invalid argument";
    case HEDFT_EGRID: return "grid edge not supported";
    case HEDFT_EHIP: return "HIP runtime error";
    case HEDFT_ENOMEM: return "device allocation failed";
    case HEDFT_ESTATE: return "call out of order";
    case HEDFT_EBOUNDS: return "impurity outside the interpolation table";
    default: return "unknown error";
  }
}

void hedft_default_params(hedft_params* p) {
  std::memset(p, 0, sizeof *p);
  p->h2o2m4 = 6.05969638298;
  p->cp4 = -2.41186e4;
  p->cpp4 = 1.85850e6;
  p->alphas = 54.31;
  p->l = 1.0;
  p->den0s = 0.04;
  p->csolid = 3.1577504e4;
  p->beta_solid = 40.0;
  p->den_m = 0.37;
  p->h4 = 2.190323;
  p->eps4 = 10.22;
  p->sigma4 = 2.556;
  p->core_otc = 0;
  p->lsolid = 0;
  p->npd = 13;
  p->icon = 13;
  p->dtmin = 1.0e-10;
  p->sdmu = 0.1;
  p->dmumax = 100.0;
  p->psimin = -5.0;
  p->lambdah = 2.0;
  for (int a = 0; a < 3; ++a) {
    p->tmean[a] = 1000.0;
    p->tsurf[a] = 2.0;
  }
  p->precision = HEDFT_PREC_FP64;
  p->lexternal = 0;
}

int hedft_grid_supported(int32_t n) {
  switch (n) {
    case 32: case 48: case 64: case 96: case 128: case 192: case 256: return 1;
    default: return 0;
  }
}

static int alloc(hedft_ctx* c, void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return HEDFT_ENOMEM;
  if (hipMemsetAsync(*p, 0, bytes, c->stream) != hipSuccess) return HEDFT_EHIP;
  return HEDFT_OK;
}

int hedft_create(const hedft_grid* g, const hedft_params* p, int32_t mode, int32_t device, hedft_ctx** out) {
  if (!g || !p || !out) return HEDFT_EINVAL;
  if (mode != HEDFT_MODE_STATICS && mode != HEDFT_MODE_TDDFT) return HEDFT_EINVAL;
  if (!hedft_grid_supported(g->nx) || !hedft_grid_supported(g->ny) || !hedft_grid_supported(g->nz))
    return HEDFT_EGRID;
  if (p->npd != 13) return HEDFT_EINVAL;
  const int icon = p->icon;
  if (!(icon == 0 || icon == 8 || icon == 13 || icon == 18 || icon == 20 || icon == 24)) return HEDFT_EINVAL;
  if (p->precision != HEDFT_PREC_FP64 && p->precision != HEDFT_PREC_MIXED) return HEDFT_EINVAL;
  // the mixed path covers real-time dynamics; imaginary-time relaxation is FP64
  if (p->precision == HEDFT_PREC_MIXED && mode != HEDFT_MODE_TDDFT) return HEDFT_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return HEDFT_EHIP;
  hedft_ctx* c = new hedft_ctx();
  std::memset((void*)c, 0, sizeof(hedft_ctx));
  c->g = *g;
  c->p = *p;
  if (c->p.lsolid) c->p.core_otc = 0;  // If(Lsolid) core4='OT ' (BCN4HeDFT.f90:654, DFT4He3d.f90:456)
  c->mode = mode;
  c->mixed = (p->precision == HEDFT_PREC_MIXED);
  c->zconv_fused = 1;
  if (const char* e = std::getenv("HEDFT_ZCONV_FUSED")) c->zconv_fused = std::atoi(e);
  // measured at 128^3 FP64: the fused stage C is 2 % slower (its 2.5-D pass
  // takes 47-50 us against 35 us for the z pass), so it is off by default
  c->fused = 0;
  if (const char* e = std::getenv("HEDFT_FUSED")) c->fused = std::atoi(e);
  // measured at 128^3: on the side stream after grad(den) for double storage,
  // on the main stream after the den convolutions for float storage
  c->lap_late = c->mixed ? 3 : 4;
  if (const char* e = std::getenv("HEDFT_LAP_LATE")) c->lap_late = std::atoi(e);
  if (const char* e = std::getenv("HEDFT_XY_SPLIT_NB")) {  // tuning override: "fwd,inv"
    c->xy_split_nb[0] = std::atoi(e);
    const char* comma = std::strchr(e, ',');
    c->xy_split_nb[1] = comma ? std::atoi(comma + 1) : c->xy_split_nb[0];
  }
  c->device = device;
  // two streams at the default priority: the main one and a side one for
  // independent branches of the mean field (joined by events, also under capture)
  if (hipStreamCreateWithFlags(&c->main_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return HEDFT_EHIP;
  }
  if (hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking) != hipSuccess) {
    hipStreamDestroy(c->main_stream);
    delete c;
    return HEDFT_EHIP;
  }
  c->stream = c->main_stream;
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  hipEventCreateWithFlags(&c->ev_mid, hipEventDisableTiming);
  Dims& d = c->d;
  d.nx = g->nx;
  d.ny = g->ny;
  d.nz = g->nz;
  d.nxh = g->nx / 2 + 1;
  d.n = (long long)d.nx * d.ny * d.nz;
  d.nh = (long long)d.nxh * d.ny * d.nz;
  c->hx = 2.0 * std::fabs(g->xmax) / d.nx;
  c->hy = 2.0 * std::fabs(g->ymax) / d.ny;
  c->hz = 2.0 * std::fabs(g->zmax) / d.nz;
  c->dxyz = c->hx * c->hy * c->hz;
  c->renor = 1.0 / ((double)d.nx * d.ny * d.nz);
  // mean-field storage: double, or float on the mixed path; psi side always double
  const size_t fw = c->mixed ? sizeof(float) : sizeof(double);
  const size_t rN = fw * d.n, cN = 2 * sizeof(double) * d.n, hN = 2 * fw * d.nh, kN = fw * d.nh;
  const size_t pN = (mode == HEDFT_MODE_TDDFT) ? cN : rN;
  int rc = HEDFT_OK;
  double** reals[] = {&c->den, &c->pot4, &c->dencg, &c->delj4, &c->denalf, &c->intx, &c->inty, &c->intz,
                      &c->ualphas, &c->dxden, &c->dyden, &c->dzden, &c->uext, &c->sto1, &c->ualph_conv,
                      &c->work1, &c->work2};
  for (double** f : reals) rc = rc ? rc : alloc(c, (void**)f, rN);
  if (mode == HEDFT_MODE_STATICS) {
    rc = rc ? rc : alloc(c, (void**)&c->timec_r, rN);
    rc = rc ? rc : alloc(c, (void**)&c->hpsi_r, rN);
    rc = rc ? rc : alloc(c, (void**)&c->psi1s, rN);
    rc = rc ? rc : alloc(c, (void**)&c->psil, rN);
    rc = rc ? rc : alloc(c, (void**)&c->psi2s, rN);
  }
  void** psis[] = {&c->psi, &c->psiold[0], &c->psiold[1], &c->psiold[2], &c->hpsiold[0], &c->hpsiold[1],
                   &c->pc, &c->q, &c->lap};
  for (void** f : psis) rc = rc ? rc : alloc(c, f, pN);
  rc = rc ? rc : alloc(c, &c->workc1, cN);
  double2** specs[] = {&c->fden, &c->spec[0], &c->spec[1], &c->spec[2], &c->spec[3], &c->workk1};
  for (double2** f : specs) rc = rc ? rc : alloc(c, (void**)f, hN);
  if (mode == HEDFT_MODE_TDDFT) rc = rc ? rc : alloc(c, (void**)&c->specd, hN);
  rc = rc ? rc : alloc(c, (void**)&c->fvlj4, kN);
  rc = rc ? rc : alloc(c, (void**)&c->wcgk, kN);
  rc = rc ? rc : alloc(c, (void**)&c->kalfs, kN);
  // partial sums: 8 regions, each large enough for the largest pencil grid x 8 values
  size_t maxb = 0;
  for (int a = 0; a < 3; ++a) {
    maxb = std::max(maxb, pencil_geom<double2>(d, a).nblocks);
    maxb = std::max(maxb, pencil_geom<double>(d, a).nblocks);
  }
  maxb = std::max(maxb, (size_t)EW_BLOCKS);
  // the 2.5-D stepper (TDDFT): z chunks of 16 planes by default, HEDFT_LAP25=<planes>
  // selects another chunk length and 0 the separate x+y Laplacian pass
  if (mode == HEDFT_MODE_TDDFT) {
    int tz = 16;
    if (const char* e = std::getenv("HEDFT_LAP25")) tz = std::atoi(e);
    if (tz > 0 && lap25_geom(d, tz).nblocks > 0) c->lap25_tz = tz;
    maxb = std::max(maxb, lap25_geom(d, 1).nblocks);
  }
  c->partials_cap = 8 * maxb * 16;  // 8 regions of 16 values per block
  rc = rc ? rc : alloc(c, (void**)&c->partials, c->partials_cap * sizeof(double));
  rc = rc ? rc : alloc(c, (void**)&c->scal, NSCAL * sizeof(double));
  rc = rc ? rc : alloc(c, (void**)&c->imp, IMP_SIZE * sizeof(double));
  rc = rc ? rc : alloc(c, (void**)&c->lstop, sizeof(int));
  rc = rc ? rc : alloc(c, (void**)&c->lamtab, sizeof(double) * (d.nx + d.ny + d.nz));
  rc = rc ? rc : alloc(c, (void**)&c->coef2_dev, 3 * sizeof(StencilCoef));
  rc = rc ? rc : alloc(c, (void**)&c->coef1_dev, 3 * sizeof(StencilCoef));
  if (rc == HEDFT_OK && hipHostMalloc((void**)&c->host_scal, NSCAL * sizeof(double)) != hipSuccess) rc = HEDFT_ENOMEM;
  if (rc == HEDFT_OK) rc = fft_setup(c);
  if (rc != HEDFT_OK) {
    hedft_destroy(c);
    return rc;
  }
  stencil_setup(c);
  if (hipMemcpyAsync(c->coef2_dev, c->coef[1], 3 * sizeof(StencilCoef), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipMemcpyAsync(c->coef1_dev, c->coef[0], 3 * sizeof(StencilCoef), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess) {
    hedft_destroy(c);
    return HEDFT_EHIP;
  }
  // spectral kernels (fforma, initcg, kalfs) and absorbing-layer tables
  std::vector<double> fv, wc, ka;
  kernel_tables(c, fv, wc, ka);
  std::vector<double> lam(d.nx + d.ny + d.nz);
  const double mx[3] = {g->xmax, g->ymax, g->zmax};
  const double hh[3] = {c->hx, c->hy, c->hz};
  const int nn[3] = {d.nx, d.ny, d.nz};
  for (int a = 0, o = 0; a < 3; o += nn[a], ++a)
    for (int i = 0; i < nn[a]; ++i)
      lam[o + i] = std::tanh((std::fabs(-mx[a] + hh[a] * i) - c->p.tmean[a]) / c->p.tsurf[a]);
  std::vector<float> fvf, wcf, kaf;
  const void *pfv = fv.data(), *pwc = wc.data(), *pka = ka.data();
  if (c->mixed) {  // tables computed in long double / double, stored rounded to FP32
    fvf.assign(fv.begin(), fv.end());
    wcf.assign(wc.begin(), wc.end());
    kaf.assign(ka.begin(), ka.end());
    pfv = fvf.data();
    pwc = wcf.data();
    pka = kaf.data();
  }
  if (hipMemcpyAsync(c->fvlj4, pfv, kN, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->wcgk, pwc, kN, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->kalfs, pka, kN, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->lamtab, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    hedft_destroy(c);
    return HEDFT_EHIP;
  }
  for (int k = 0; k < 3; ++k) c->ioldp[k] = c->ioldr[k] = c->ioldv[k] = k;
  for (int k = 0; k < 2; ++k) c->ioldh[k] = c->iolda[k] = k;
  *out = c;
  return HEDFT_OK;
}

int hedft_destroy(hedft_ctx* c) {
  if (!c) return HEDFT_OK;
  hipSetDevice(c->device);
  if (c->side_stream) hipStreamSynchronize(c->side_stream);
  if (c->main_stream) hipStreamSynchronize(c->main_stream);
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  void* ptrs[] = {c->den, c->pot4, c->dencg, c->delj4, c->denalf, c->intx, c->inty, c->intz, c->ualphas,
                  c->dxden, c->dyden, c->dzden, c->uext, c->sto1, c->ualph_conv, c->work1, c->work2,
                  c->timec_r, c->hpsi_r, c->psi1s, c->psi2s, c->psil, c->psi, c->psiold[0], c->psiold[1], c->psiold[2], c->hpsiold[0],
                  c->hpsiold[1], c->pc, c->q, c->lap, c->workc1, c->fden, c->spec[0], c->spec[1], c->spec[2],
                  c->spec[3], c->workk1, c->fvlj4, c->wcgk, c->kalfs, c->tw, c->twf, c->lamtab, c->partials, c->scal,
                  c->imp, c->lstop, c->potion, c->psix, c->psixold[0], c->psixold[1], c->psixold[2],
                  c->hpsixold[0], c->hpsixold[1], c->pcx, c->qx, c->lapx, c->potx4, c->upotx, c->uimpk,
                  c->specq[0], c->specq[1], c->specd, c->coef2_dev, c->coef1_dev};
  for (void* q : ptrs)
    if (q) hipFree(q);
  clear_pairs(c);
  if (c->host_scal) hipHostFree(c->host_scal);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->ev_mid) hipEventDestroy(c->ev_mid);
  if (c->side_stream) hipStreamDestroy(c->side_stream);
  if (c->main_stream) hipStreamDestroy(c->main_stream);
  delete c;
  return HEDFT_OK;
}

// device pointer, byte size and whether the field is computed on download
// bytes: the host-side (reference layout, FP64) size.  On the mixed path every
// field except psi, its PC history and the computed ones is stored as float on
// the device; *low reports that (uploads and downloads then convert).
static int field_ptr(hedft_ctx* c, int f, void** ptr, size_t* bytes, int* computed, int* low = nullptr) {
  const size_t rN = sizeof(double) * c->d.n, cN = 2 * rN, hN = sizeof(double2) * c->d.nh,
               kN = sizeof(double) * c->d.nh;
  const bool td = (c->mode == HEDFT_MODE_TDDFT);
  const size_t pN = td ? cN : rN;
  *computed = 0;
  switch (f) {
    case HEDFT_DEN: *ptr = c->den; *bytes = rN; break;
    case HEDFT_PSI: *ptr = c->psi; *bytes = pN; break;
    case HEDFT_POT4: *ptr = c->pot4; *bytes = rN; break;
    case HEDFT_DENCG: *ptr = c->dencg; *bytes = rN; break;
    case HEDFT_DELJ4: *ptr = c->delj4; *bytes = rN; break;
    case HEDFT_DENALF: *ptr = c->denalf; *bytes = rN; break;
    case HEDFT_FALFS: *ptr = c->work2; *bytes = rN; *computed = 1; break;
    case HEDFT_INTXALF: *ptr = c->intx; *bytes = rN; break;
    case HEDFT_INTYALF: *ptr = c->inty; *bytes = rN; break;
    case HEDFT_INTZALF: *ptr = c->intz; *bytes = rN; break;
    case HEDFT_UALPHAS: *ptr = c->ualphas; *bytes = rN; break;
    case HEDFT_DXDEN: *ptr = c->dxden; *bytes = rN; break;
    case HEDFT_DYDEN: *ptr = c->dyden; *bytes = rN; break;
    case HEDFT_DZDEN: *ptr = c->dzden; *bytes = rN; break;
    case HEDFT_UIMP: *ptr = c->work2; *bytes = rN; *computed = 1; break;
    case HEDFT_UEXT: *ptr = c->uext; *bytes = rN; break;
    case HEDFT_TIMEC:
      if (td) { *ptr = c->workc1; *bytes = cN; *computed = 1; }
      else { *ptr = c->timec_r; *bytes = rN; }
      break;
    case HEDFT_PSIOLD1: case HEDFT_PSIOLD2: case HEDFT_PSIOLD3:
      *ptr = c->psiold[f - HEDFT_PSIOLD1]; *bytes = pN; break;
    case HEDFT_HPSIOLD1: case HEDFT_HPSIOLD2:
      *ptr = c->hpsiold[f - HEDFT_HPSIOLD1]; *bytes = pN; break;
    case HEDFT_PC: *ptr = c->pc; *bytes = pN; break;
    case HEDFT_Q: *ptr = c->q; *bytes = pN; break;
    case HEDFT_FDEN: *ptr = c->fden; *bytes = hN; break;
    case HEDFT_FVLJ4: *ptr = c->fvlj4; *bytes = kN; break;
    case HEDFT_WCGK: *ptr = c->wcgk; *bytes = kN; break;
    case HEDFT_KALFS: *ptr = c->kalfs; *bytes = kN; break;
    case HEDFT_WORK1: *ptr = c->work1; *bytes = rN; break;
    case HEDFT_WORK2: *ptr = c->work2; *bytes = rN; break;
    case HEDFT_WORKC1: *ptr = c->workc1; *bytes = cN; break;
    case HEDFT_WORKK1: *ptr = c->workk1; *bytes = hN; break;
    case HEDFT_PSIX: case HEDFT_POTX4: case HEDFT_UPOTX: case HEDFT_DENX: case HEDFT_UIMPK:
    case HEDFT_PSIXOLD1: case HEDFT_PSIXOLD2: case HEDFT_PSIXOLD3: case HEDFT_HPSIXOLD1:
    case HEDFT_HPSIXOLD2: case HEDFT_PCX: case HEDFT_QX:
      if (!c->quantum) return HEDFT_ESTATE;
      if (!td) {  // statics: real psix, its potentials and kernel
        switch (f) {
          case HEDFT_PSIX: *ptr = c->psix; *bytes = rN; break;
          case HEDFT_POTX4: *ptr = c->potx4; *bytes = rN; break;
          case HEDFT_UPOTX: *ptr = c->upotx; *bytes = rN; break;
          case HEDFT_DENX: *ptr = c->workc1; *bytes = rN; *computed = 1; break;
          case HEDFT_UIMPK: *ptr = c->uimpk; *bytes = kN; break;
          default: return HEDFT_EINVAL;
        }
        break;
      }
      switch (f) {
        case HEDFT_PSIX: *ptr = c->psix; *bytes = cN; break;
        case HEDFT_POTX4: *ptr = c->potx4; *bytes = rN; break;
        case HEDFT_UPOTX: *ptr = c->upotx; *bytes = rN; break;
        case HEDFT_DENX: *ptr = c->workc1; *bytes = rN; *computed = 1; break;
        case HEDFT_UIMPK: *ptr = c->uimpk; *bytes = kN; break;
        case HEDFT_PCX: *ptr = c->pcx; *bytes = cN; break;
        case HEDFT_QX: *ptr = c->qx; *bytes = cN; break;
        case HEDFT_HPSIXOLD1: case HEDFT_HPSIXOLD2: *ptr = c->hpsixold[f - HEDFT_HPSIXOLD1]; *bytes = cN; break;
        default: *ptr = c->psixold[f - HEDFT_PSIXOLD1]; *bytes = cN; break;
      }
      break;
    case HEDFT_PSIL:
      if (td) return HEDFT_EINVAL;
      *ptr = c->psil; *bytes = rN; break;
    default: return HEDFT_EINVAL;
  }
  if (*computed && f != HEDFT_TIMEC) *ptr = c->workc1;  // evaluated into double scratch
  if (low) {
    const bool psi_side = (f == HEDFT_PSI || (f >= HEDFT_PSIOLD1 && f <= HEDFT_Q) || f == HEDFT_WORKC1 ||
                           (f == HEDFT_TIMEC && td) || f == HEDFT_PSIX || (f >= HEDFT_PSIXOLD1 && f <= HEDFT_QX));
    *low = c->mixed && !*computed && !psi_side;
  }
  return HEDFT_OK;
}

size_t hedft_field_bytes(const hedft_ctx* c, int32_t field) {
  void* p;
  size_t b = 0;
  int comp;
  if (field_ptr(const_cast<hedft_ctx*>(c), field, &p, &b, &comp) != HEDFT_OK) return 0;
  return b;
}

int hedft_upload(hedft_ctx* c, int32_t field, const void* host, size_t bytes) {
  if (!c || !host) return HEDFT_EINVAL;
  void* p;
  size_t b;
  int comp;
  int low = 0;
  HEDFT_TRY(field_ptr(c, field, &p, &b, &comp, &low));
  if (comp || bytes != b) return HEDFT_EINVAL;
  hipSetDevice(c->device);
  if (low) {  // FP64 host data -> double scratch -> float field
    HEDFT_CHECK(hipMemcpyAsync(c->workc1, host, b, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, c->stream, (long long)(b / 8),
                       (const double*)c->workc1, (float*)p);
  } else {
    HEDFT_CHECK(hipMemcpyAsync(p, host, b, hipMemcpyHostToDevice, c->stream));
  }
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  c->graph_valid = 0;
  c->lap_fresh = 0;
  return HEDFT_OK;
}

int hedft_download(hedft_ctx* c, int32_t field, void* host, size_t bytes) {
  if (!c || !host) return HEDFT_EINVAL;
  void* p;
  size_t b;
  int comp;
  int low = 0;
  HEDFT_TRY(field_ptr(c, field, &p, &b, &comp, &low));
  if (bytes != b) return HEDFT_EINVAL;
  hipSetDevice(c->device);
  const dim3 gr(1024), bl(256);
  if (comp) {
    double* w = (double*)c->workc1;
    if (field == HEDFT_UIMP) hipLaunchKernelGGL(k_eval_uimp, gr, bl, 0, c->stream, c->d, make_impdev(c), w);
    else if (field == HEDFT_DENX && c->mode == HEDFT_MODE_STATICS)
      hipLaunchKernelGGL(k_sq, gr, bl, 0, c->stream, c->d.n, rz(c->psix), w);
    else if (field == HEDFT_DENX)
      hipLaunchKernelGGL(k_abs2, gr, bl, 0, c->stream, c->d.n, (const double2*)c->psix, w);
    else if (field == HEDFT_FALFS && c->mixed)
      hipLaunchKernelGGL(k_eval_falfs<float>, gr, bl, 0, c->stream, c->d.n, (const float*)c->denalf,
                         c->p.den0s, w);
    else if (field == HEDFT_FALFS)
      hipLaunchKernelGGL(k_eval_falfs<double>, gr, bl, 0, c->stream, c->d.n, c->denalf, c->p.den0s, w);
    else hipLaunchKernelGGL(k_eval_timec, gr, bl, 0, c->stream, c->d, make_lam(c), (double2*)c->workc1);
  }
  if (low) {  // float field -> double scratch -> FP64 host data
    hipLaunchKernelGGL((k_convert<float, double>), gr, bl, 0, c->stream, (long long)(b / 8), (const float*)p,
                       (double*)c->workc1);
    p = c->workc1;
  }
  HEDFT_CHECK(hipMemcpyAsync(host, p, b, hipMemcpyDeviceToHost, c->stream));
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  return HEDFT_OK;
}

int hedft_synchronize(hedft_ctx* c) {
  if (!c) return HEDFT_EINVAL;
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  return HEDFT_OK;
}

static int done(hedft_ctx* c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "hedft: %s\n", hipGetErrorString(e));
    return HEDFT_EHIP;
  }
  (void)c;
  return HEDFT_OK;
}

int hedft_derivnD(hedft_ctx* c, int32_t id, int32_t iv, int32_t in, int32_t out) {
  if (!c || id < 1 || id > 2 || iv < 1 || iv > 3 || c->mixed) return HEDFT_EINVAL;
  void *pi, *po;
  size_t bi, bo;
  int ci, co;
  HEDFT_TRY(field_ptr(c, in, &pi, &bi, &ci));
  HEDFT_TRY(field_ptr(c, out, &po, &bo, &co));
  if (ci || co || bi != bo || pi == po) return HEDFT_EINVAL;
  const size_t rN = sizeof(double) * c->d.n;
  if (bi != rN && bi != 2 * rN) return HEDFT_EINVAL;
  c->lap_fresh = 0;
  HEDFT_TRY(deriv_plain(c, id, iv - 1, pi, po, bi == 2 * rN));
  return done(c);
}

int hedft_fftfw(hedft_ctx* c, int32_t in, int32_t out) {
  if (!c || c->mixed) return HEDFT_EINVAL;  // FP64 parity hook
  void *pi, *po;
  size_t bi, bo;
  int ci, co;
  HEDFT_TRY(field_ptr(c, in, &pi, &bi, &ci));
  HEDFT_TRY(field_ptr(c, out, &po, &bo, &co));
  if (ci || co || bi != sizeof(double) * c->d.n || bo != sizeof(double2) * c->d.nh) return HEDFT_EINVAL;
  HEDFT_TRY(fftfw(c, (const double*)pi, (double2*)po));
  return done(c);
}

int hedft_fftbk(hedft_ctx* c, int32_t in, int32_t out) {
  if (!c || c->mixed) return HEDFT_EINVAL;  // FP64 parity hook
  void *pi, *po;
  size_t bi, bo;
  int ci, co;
  HEDFT_TRY(field_ptr(c, in, &pi, &bi, &ci));
  HEDFT_TRY(field_ptr(c, out, &po, &bo, &co));
  if (ci || co || bo != sizeof(double) * c->d.n || bi != sizeof(double2) * c->d.nh) return HEDFT_EINVAL;
  c->lap_fresh = 0;
  HEDFT_TRY(fftbk(c, (double2*)pi, (double*)po));
  return done(c);
}

int hedft_derden(hedft_ctx* c) {
  HEDFT_TRY(derden_force(c, 0, 0));
  return done(c);
}

int hedft_poten(hedft_ctx* c) {
  HEDFT_TRY(poten(c, 0));
  return done(c);
}

int hedft_term_alfa(hedft_ctx* c) {
  if (!c->p.core_otc) return HEDFT_OK;
  // pot4 is re-assembled by the divergence pass from the stored delj4/dencg/sto1;
  // keep it untouched for a standalone call by saving and restoring it
  const size_t b = (c->mixed ? sizeof(float) : sizeof(double)) * c->d.n;
  HEDFT_CHECK(hipMemcpyAsync(c->work2, c->pot4, b, hipMemcpyDeviceToDevice, c->stream));
  HEDFT_TRY(c->mixed ? alpha_term<float>(c) : alpha_term<double>(c));
  HEDFT_CHECK(hipMemcpyAsync(c->pot4, c->work2, b, hipMemcpyDeviceToDevice, c->stream));
  return done(c);
}

int hedft_energy(hedft_ctx* c, hedft_energies* e) {
  if (!c || !e) return HEDFT_EINVAL;
  double s[8];
  if (c->mode == HEDFT_MODE_TDDFT) {
    if (c->mixed) HEDFT_TRY((energy_sums<double2, float>(c, s)));
    else HEDFT_TRY((energy_sums<double2, double>(c, s)));
  } else {
    HEDFT_TRY((energy_sums<double, double>(c, s)));
  }
  const hedft_params& p = c->p;
  const double dv = c->dxyz;
  std::memset(e, 0, sizeof *e);
  e->ekin4 = p.h2o2m4 * s[0] * dv;
  e->elj4 = 0.5 * s[1] * dv;
  e->ecor4 = s[2] * dv;
  e->ealphas = p.core_otc ? -p.h2o2m4 * 0.5 * p.alphas * s[3] * dv : 0.0;
  e->esolid = p.lsolid ? p.csolid * s[4] * dv : 0.0;
  e->etot4 = e->ekin4 + e->elj4 + e->ecor4 + e->esolid + e->ealphas;
  if (c->mode == HEDFT_MODE_TDDFT) {
    e->eext = s[5] * dv;
    e->etot4 += e->eext;
    double rv[12];
    HEDFT_CHECK(hipMemcpy(rv, c->imp + IMP_V, sizeof rv, hipMemcpyDeviceToHost));
    for (int k = 0; k < c->nimp; ++k) {
      double v2 = 0.0;
      for (int a = 0; a < 3; ++a) v2 += rv[k + c->nimp * a] * rv[k + c->nimp * a];
      e->ekinx += 0.5 * c->mass[k] * v2;
    }
    e->eHeX = s[6] * dv;
    if (c->nimp >= 2 && c->pair_v_host) {  // energy.f90:168-177
      double r[3 * HEDFT_MAX_IMP];
      HEDFT_CHECK(hipMemcpy(r, c->imp + IMP_R, sizeof(double) * 3 * c->nimp, hipMemcpyDeviceToHost));
      double eij = 0.0;
      const int ni = c->nimp;
      for (int k = 0; k < ni; ++k)
        for (int m = k + 1; m < ni; ++m) {
          const double dx = r[k] - r[m], dy = r[k + ni] - r[m + ni], dz = r[k + 2 * ni] - r[m + 2 * ni];
          const double aux1 = std::sqrt(dx * dx + dy * dy + dz * dz);
          eij += radial_lerp(c->pair_v_host + (size_t)pair_index(k, m, ni) * c->npair, c->npair, c->delpair, aux1,
                             nullptr);
        }
      e->eimpu_impu = eij;
    }
    if (c->quantum) {
      double sx[2];
      if (c->mixed) HEDFT_TRY(energy_sums_x<float>(c, sx));
      else HEDFT_TRY(energy_sums_x<double>(c, sx));
      e->ekinx = c->h2o2mimp * sx[0] * dv;
      e->eHeX = sx[1] * dv;
    }
    e->eimpu = e->ekinx + e->eHeX + e->eimpu_impu;
  } else if (c->quantum) {
    // He_statics/energy.f90:76-93,137-149: etot = etot4 + ekinx + eimpu
    double sk;
    HEDFT_TRY(energy_sums_xs(c, &sk));
    e->ekinx = c->h2o2mimp * sk * dv;
    e->eimpu = s[5] * dv;
    e->etot = e->etot4 + e->ekinx + e->eimpu;
    e->norm = s[7] * dv;
    return HEDFT_OK;
  } else {
    e->eext = p.lexternal ? s[5] * dv : 0.0;
    e->eimpu = e->eext;
  }
  e->etot = e->etot4 + e->eimpu;
  e->norm = s[7] * dv;
  return HEDFT_OK;
}

int hedft_compute_observables(hedft_ctx* c, int32_t n4, double mhe, hedft_observables* o) {
  if (!c || !o || c->mode != HEDFT_MODE_TDDFT || n4 <= 0 || !(mhe > 0.0)) return HEDFT_EINVAL;
  double s[13];
  if (c->mixed) HEDFT_TRY(observable_sums<float>(c, s));
  else HEDFT_TRY(observable_sums<double>(c, s));
  const double dv = c->dxyz;
  std::memset(o, 0, sizeof *o);
  o->norm = s[0] * dv;
  for (int a = 0; a < 3; ++a) {
    o->rcm[a] = s[1 + a] * dv / (double)n4;
    o->r2[a] = s[4 + a] * dv;
    o->vcom[a] = -s[10 + a] * dv / (o->norm * mhe);
  }
  const double xc = o->rcm[0], yc = o->rcm[1], zc = o->rcm[2];
  const double gx = s[10], gy = s[11], gz = s[12];
  o->l[0] = (s[7] - zc * gy + yc * gz) * dv;
  o->l[1] = (s[8] - xc * gz + zc * gx) * dv;
  o->l[2] = (s[9] - yc * gx + xc * gy) * dv;
  return HEDFT_OK;
}

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

int hedft_steprk(hedft_ctx* c, double deltat) {
  if (!c || c->mode != HEDFT_MODE_TDDFT) return HEDFT_EINVAL;
  HEDFT_TRY(imp_ready(c));
  HEDFT_TRY(steprk_impl(c, deltat));
  return done(c);
}

static int read_pc_errors(hedft_ctx* c, double* errHe, double* errimp, double* errvimp) {
  if (!errHe && !errimp && !errvimp) return HEDFT_OK;
  double s[8];
  HEDFT_TRY(read_scal(c, 0, 8, s));
  if (errHe) *errHe = (9.0 / 121.0) * s[S_ERRPC] / (double)c->d.n;
  if (errimp) *errimp = c->quantum ? (9.0 / 121.0) * s[S_ERRPCX] / (double)c->d.n : c->nimp ? s[S_ERRIMP] : 0.0;
  if (errvimp) *errvimp = c->nimp ? s[S_ERRVIMP] : 0.0;
  return HEDFT_OK;
}

int hedft_steppc(hedft_ctx* c, double deltat, double* errHe, double* errimp, double* errvimp) {
  if (!c || c->mode != HEDFT_MODE_TDDFT) return HEDFT_EINVAL;
  HEDFT_TRY(imp_ready(c));
  if (!c->rk_done) return HEDFT_ESTATE;
  HEDFT_TRY(steppc_impl(c, deltat));
  flush_side_pre(c);  // STEPPC ends with the impurity corrector and the errHe sum
  HEDFT_TRY(read_pc_errors(c, errHe, errimp, errvimp));
  return done(c);
}

int hedft_set_impurities(hedft_ctx* c, int32_t nimp, int32_t npot, double delinter, double rmaxinterpol,
                         const double* potion, const double* mass) {
  if (!c || nimp < 0 || nimp > HEDFT_MAX_IMP || (nimp > 0 && (!potion || !mass || npot < 2))) return HEDFT_EINVAL;
  if (nimp > 0 && c->quantum) return HEDFT_EINVAL;  // one impurity model per context
  HEDFT_CHECK(hipStreamSynchronize(c->stream));  // queued work may still read the old table
  // a captured graph holds the old impurity count and table pointer
  c->graph_valid = 0;
  if (c->potion) hipFree(c->potion);
  c->potion = nullptr;
  clear_pairs(c);  // pair tables belong to one impurity set
  c->nimp = 0;
  c->npot = npot;
  c->delinter = delinter;
  c->rmaxinterpol = rmaxinterpol;
  if (nimp == 0) return HEDFT_OK;
  if (hipMalloc(&c->potion, sizeof(double) * (size_t)nimp * npot) != hipSuccess) {
    c->potion = nullptr;
    return HEDFT_ENOMEM;
  }
  HEDFT_CHECK(hipMemcpy(c->potion, potion, sizeof(double) * (size_t)nimp * npot, hipMemcpyHostToDevice));
  c->nimp = nimp;
  for (int k = 0; k < nimp; ++k) c->mass[k] = mass[k];
  double m[HEDFT_MAX_IMP] = {0};
  for (int k = 0; k < nimp; ++k) m[k] = mass[k];
  HEDFT_CHECK(hipMemcpy(c->imp + IMP_MASS, m, sizeof m, hipMemcpyHostToDevice));
  c->graph_valid = 0;
  return HEDFT_OK;
}

int hedft_set_impurity_state(hedft_ctx* c, const double* rimp, const double* vimp, const double* aimp) {
  if (!c) return HEDFT_EINVAL;
  const size_t b = sizeof(double) * 3 * c->nimp;
  // ordered after the work already queued on the context's stream
  if (rimp) HEDFT_CHECK(hipMemcpyAsync(c->imp + IMP_R, rimp, b, hipMemcpyHostToDevice, c->stream));
  if (vimp) HEDFT_CHECK(hipMemcpyAsync(c->imp + IMP_V, vimp, b, hipMemcpyHostToDevice, c->stream));
  if (aimp) HEDFT_CHECK(hipMemcpyAsync(c->imp + IMP_A, aimp, b, hipMemcpyHostToDevice, c->stream));
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  return HEDFT_OK;
}

int hedft_get_impurity_state(hedft_ctx* c, double* rimp, double* vimp, double* aimp, double* force,
                             int32_t* lstopimp) {
  if (!c) return HEDFT_EINVAL;
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  const size_t b = sizeof(double) * 3 * c->nimp;
  if (rimp) HEDFT_CHECK(hipMemcpy(rimp, c->imp + IMP_R, b, hipMemcpyDeviceToHost));
  if (vimp) HEDFT_CHECK(hipMemcpy(vimp, c->imp + IMP_V, b, hipMemcpyDeviceToHost));
  if (aimp) HEDFT_CHECK(hipMemcpy(aimp, c->imp + IMP_A, b, hipMemcpyDeviceToHost));
  if (force) HEDFT_CHECK(hipMemcpy(force, c->imp + IMP_F, b, hipMemcpyDeviceToHost));
  if (lstopimp) {
    int v = 0;
    HEDFT_CHECK(hipMemcpy(&v, c->lstop, sizeof v, hipMemcpyDeviceToHost));
    *lstopimp = v;
  }
  return HEDFT_OK;
}

__global__ void k_check_bounds(ImpDev im, Dims d) {
  // the farthest grid point from each impurity is a box corner
  if (threadIdx.x != 0) return;
  for (int k = 0; k < im.nimp; ++k)
    for (int cz_ = 0; cz_ < 2; ++cz_)
      for (int cy = 0; cy < 2; ++cy)
        for (int cx = 0; cx < 2; ++cx) {
          double xx, yy, zz;
          imp_xyz(im, cx ? d.nx - 1 : 0, cy ? d.ny - 1 : 0, cz_ ? d.nz - 1 : 0, xx, yy, zz);
          imp_u_k(im, k, xx, yy, zz, true);
        }
}

int hedft_potenimp(hedft_ctx* c) {
  if (!c) return HEDFT_EINVAL;
  if (c->nimp > 0) hipLaunchKernelGGL(k_check_bounds, dim3(1), dim3(64), 0, c->stream, make_impdev(c), c->d);
  if (c->quantum) HEDFT_TRY(c->mixed ? potenimp_q<float>(c) : potenimp_q<double>(c));
  return done(c);
}

int hedft_update_aimp(hedft_ctx* c) {
  if (!c) return HEDFT_EINVAL;
  if (c->nimp > 0) hipLaunchKernelGGL(k_imp_aimp, dim3(1), dim3(1), 0, c->stream, c->imp, c->nimp);
  return done(c);
}

// potenimpini's uimp_k(:,:,:,1) and the mass of one quantum impurity
// (Dynamics_quantum_impurities/potenimpini.f90:19-29, steprk.f90:86)
// He_statics with limp: real psix, psi1x, psi2x, the kernel vq (BCN4HeDFT.f90:738-791)
// and h2o2mx = h2o2m4 (4.002602/mimpur) (BCN4HeDFT.f90:537)
static int set_quantum_statics(hedft_ctx* c, const double* vq, size_t bytes, double m_imp_u) {
  if (c->p.lexternal) return HEDFT_EINVAL;  // one external field in the pot4 assembly
  const size_t rN = sizeof(double) * c->d.n, kN = sizeof(double) * c->d.nh, hN = sizeof(double2) * c->d.nh;
  int rc = HEDFT_OK;
  void** reals[] = {&c->psix, &c->psixold[0], &c->psixold[1], (void**)&c->potx4, (void**)&c->upotx};
  for (void** f : reals)
    if (!*f) rc = rc ? rc : alloc(c, f, rN);
  if (!c->uimpk) rc = rc ? rc : alloc(c, (void**)&c->uimpk, kN);
  if (!c->specq[0]) rc = rc ? rc : alloc(c, (void**)&c->specq[0], hN);
  if (rc != HEDFT_OK) return rc;
  HEDFT_CHECK(hipMemcpyAsync(c->uimpk, vq, bytes, hipMemcpyHostToDevice, c->stream));
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  c->m_imp_u = m_imp_u;
  c->h2o2mimp = c->p.h2o2m4 * (4.002602 / m_imp_u);
  c->mu4 = 0.0;
  c->quantum = 1;
  return done(c);
}

int hedft_set_quantum_impurity(hedft_ctx* c, const double* uimp_k, size_t bytes, double m_imp_u, double mu4) {
  if (!c || !uimp_k || c->nimp > 0 || !(m_imp_u > 0.0)) return HEDFT_EINVAL;
  if (bytes != sizeof(double) * (size_t)c->d.nh) return HEDFT_EINVAL;
  hipSetDevice(c->device);
  HEDFT_CHECK(hipStreamSynchronize(c->stream));  // queued work may still read the old kernel
  if (c->mode == HEDFT_MODE_STATICS) return set_quantum_statics(c, uimp_k, bytes, m_imp_u);
  const size_t fw = c->mixed ? sizeof(float) : sizeof(double);
  const size_t cN = sizeof(double2) * c->d.n, rN = fw * c->d.n, kN = fw * c->d.nh, hN = 2 * fw * c->d.nh;
  int rc = HEDFT_OK;
  void** cplx[] = {&c->psix, &c->psixold[0], &c->psixold[1], &c->psixold[2], &c->hpsixold[0], &c->hpsixold[1],
                   &c->pcx, &c->qx, &c->lapx};
  for (void** f : cplx)
    if (!*f) rc = rc ? rc : alloc(c, f, cN);
  if (!c->potx4) rc = rc ? rc : alloc(c, (void**)&c->potx4, rN);
  if (!c->upotx) rc = rc ? rc : alloc(c, (void**)&c->upotx, rN);
  if (!c->uimpk) rc = rc ? rc : alloc(c, (void**)&c->uimpk, kN);
  for (int k = 0; k < 2; ++k)
    if (!c->specq[k]) rc = rc ? rc : alloc(c, (void**)&c->specq[k], hN);
  if (rc != HEDFT_OK) return rc;
  if (c->mixed) {  // FP64 table -> double scratch -> float kernel
    HEDFT_CHECK(hipMemcpyAsync(c->workc1, uimp_k, bytes, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, c->stream, (long long)c->d.nh,
                       (const double*)c->workc1, (float*)c->uimpk);
  } else {
    HEDFT_CHECK(hipMemcpyAsync(c->uimpk, uimp_k, bytes, hipMemcpyHostToDevice, c->stream));
  }
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  c->m_imp_u = m_imp_u;
  c->h2o2mimp = c->p.h2o2m4 * 4.002603 / m_imp_u;
  c->mu4 = mu4;
  c->quantum = 1;
  c->graph_valid = 0;
  c->lap_fresh = 0;
  return done(c);
}

int hedft_quantum_centroid(hedft_ctx* c, double* norm, double* rcm) {
  if (!c || !c->quantum) return HEDFT_EINVAL;
  hipLaunchKernelGGL(k_centroid, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, (const double2*)c->psix,
                     -c->g.xmax, -c->g.ymax, -c->g.zmax, c->hx, c->hy, c->hz, gen_partials(c));
  hipLaunchKernelGGL(k_finalize<4>, dim3(1), dim3(256), 0, c->stream, gen_partials(c), EW_BLOCKS,
                     c->scal + S_ENERGYQ, 1.0);
  double s[4];
  HEDFT_TRY(read_scal(c, S_ENERGYQ, 4, s));
  const double nrm = s[0] * c->dxyz;
  if (norm) *norm = nrm;
  if (rcm)
    for (int a = 0; a < 3; ++a) rcm[a] = s[1 + a] * c->dxyz / nrm;
  return HEDFT_OK;
}

int hedft_statics_init(hedft_ctx* c, double n4, double* mu4) {
  if (!c || c->mode != HEDFT_MODE_STATICS) return HEDFT_EINVAL;
  hipLaunchKernelGGL(k_sum1, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->den, gen_partials(c));
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  hipLaunchKernelGGL(k_statics_psi0, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->den, rz(c->psi), c->scal,
                     n4, c->dxyz);
  if (c->quantum) {  // psi1x = psi2x = psix (BCN4HeDFT.f90:781-785)
    const size_t b = sizeof(double) * c->d.n;
    HEDFT_CHECK(hipMemcpyAsync(c->psixold[0], c->psix, b, hipMemcpyDeviceToDevice, c->stream));
    HEDFT_CHECK(hipMemcpyAsync(c->psixold[1], c->psix, b, hipMemcpyDeviceToDevice, c->stream));
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
  HEDFT_TRY(lap_xy(c, c->psix, 0));
  OutHpsiR o;
  o.partials = gen_partials(c);
  o.lapxy = rz(c->lap);
  o.psi = rz(c->psix);
  o.pot4 = c->upotx;
  o.hpsi = c->hpsi_r;
  o.h2o2m4 = c->h2o2mimp;
  launch_pencil<double, 2>(c->stream, c->d, c->coef[1][2], InPlain<double>{rz(c->psix)}, o);
  finalize<1>(c, gen_partials(c), pencil_geom<double>(c->d, 2).nblocks, S_MU, c->dxyz);
  const double eold = *epsx;
  hipLaunchKernelGGL(k_evolox_g, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->hpsi_r, rz(c->psix), eold);
  double* g = c->hpsi_r;
  for (int k = 0; k < ironx; ++k) {  // ping-pong through work1 / work2
    double* dst = (g == c->work1) ? c->work2 : c->work1;
    hipLaunchKernelGGL(k_ironing, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, g, dst);
    g = dst;
  }
  hipLaunchKernelGGL(k_evolox_new, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psix),
                     rz(c->psixold[0]), rz(c->psixold[1]), g, deltatx, vdt1, vdt2, c->p.psimin, gen_partials(c));
  finalize<1>(c, gen_partials(c), EW_BLOCKS, S_SUMSQ, 1.0);
  hipLaunchKernelGGL(k_norm1, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, rz(c->psix), c->scal, c->dxyz);
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

int hedft_tddft_init(hedft_ctx* c) {
  if (!c || c->mode != HEDFT_MODE_TDDFT) return HEDFT_EINVAL;
  HEDFT_TRY(imp_ready(c));
  c->lap_fresh = 0;
  if (c->mixed)
    hipLaunchKernelGGL(k_tddft_psi0<float>, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, fld<float>(c->den),
                       (double2*)c->psi);
  else
    hipLaunchKernelGGL(k_tddft_psi0<double>, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, c->den,
                       (double2*)c->psi);
  // qr = qv = rimpold = vimpold = aimpold = 0 (DFT4He3d.f90:625-629)
  HEDFT_CHECK(hipMemsetAsync(c->imp + IMP_QR, 0, sizeof(double) * 24, c->stream));
  HEDFT_CHECK(hipMemsetAsync(c->imp + IMP_ROLD, 0, sizeof(double) * (IMP_MASS - IMP_ROLD), c->stream));
  // qx = 0 (Dynamics_quantum_impurities/DFT4He3d.f90:617)
  if (c->quantum) HEDFT_CHECK(hipMemsetAsync(c->qx, 0, sizeof(double2) * c->d.n, c->stream));
  HEDFT_TRY(hedft_potenimp(c));
  HEDFT_TRY(tddft_after_step(c));
  c->rk_done = 0;
  return done(c);
}

int hedft_tddft_iteration(hedft_ctx* c, double deltat, int32_t iter, int32_t nrk, double* errHe, double* errimp,
                          double* errvimp) {
  if (!c || c->mode != HEDFT_MODE_TDDFT) return HEDFT_EINVAL;
  HEDFT_TRY(imp_ready(c));
  if (iter <= nrk) {
    HEDFT_TRY(steprk_impl(c, deltat));
    if (errHe) *errHe = 0.0;
    if (errimp) *errimp = 0.0;
    if (errvimp) *errvimp = 0.0;
    HEDFT_TRY(tddft_after_step(c));
  } else {
    if (!c->rk_done) return HEDFT_ESTATE;
    HEDFT_TRY(steppc_impl(c, deltat));
    HEDFT_TRY(tddft_after_step(c));
    HEDFT_TRY(read_pc_errors(c, errHe, errimp, errvimp));
  }
  return done(c);
}

// One PC iteration per slot phase; the slot rotation has period 6, so six
// consecutive iterations are captured into one graph and replayed.
}  // extern "C"

namespace hedft {
static constexpr int PC_PERIOD = 6;  // slot rotation period of the PC history

static void slot_state(const hedft_ctx* c, int* s) {
  std::memcpy(s, c->ioldp, 3 * sizeof(int));
  std::memcpy(s + 3, c->ioldh, 2 * sizeof(int));
  std::memcpy(s + 5, c->ioldr, 3 * sizeof(int));
  std::memcpy(s + 8, c->ioldv, 3 * sizeof(int));
  std::memcpy(s + 11, c->iolda, 2 * sizeof(int));
}

// Prepares a context for graph replay: x+y Laplacian of psi current, and a
// graph of PC_PERIOD iterations captured from the current slot state and dt.
static int ensure_graph(hedft_ctx* c, double deltat) {
  // the captured iterations start from a prepared x+y Laplacian of psi (as every
  // iteration leaves it), so prepare it now if an upload invalidated it
  HEDFT_TRY(prepare_lap(c));
  c->lap_fresh = 1;
  int slots[13];
  slot_state(c, slots);
  if (c->graph_valid && c->graph_dt == deltat && std::memcmp(slots, c->graph_slots, sizeof slots) == 0)
    return HEDFT_OK;
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_valid = 0;
  HEDFT_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = HEDFT_OK;
  for (int s = 0; s < PC_PERIOD && rc == HEDFT_OK; ++s) {
    rc = steppc_impl(c, deltat);
    if (rc == HEDFT_OK) rc = tddft_after_step(c);
  }
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &gr);
  if (rc != HEDFT_OK) {
    if (gr) hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return HEDFT_EHIP;
  c->graph = gr;
  HEDFT_CHECK(hipGraphInstantiate(&c->graph_exec, gr, nullptr, nullptr, 0));
  // capture advanced the host-side slot indices by one full period, i.e. back
  // to where they started
  c->graph_valid = 1;
  c->graph_dt = deltat;
  std::memcpy(c->graph_slots, slots, sizeof slots);
  return HEDFT_OK;
}

// enqueue nsteps PC iterations: graph replays for whole periods, eager steps for the rest
static int enqueue_pc(hedft_ctx* c, double deltat, int nsteps, int use_graph) {
  int done_steps = 0;
  use_graph = use_graph && !c->prefer_eager;
  if (use_graph && nsteps >= PC_PERIOD) {
    HEDFT_TRY(ensure_graph(c, deltat));
    for (; done_steps + PC_PERIOD <= nsteps; done_steps += PC_PERIOD)
      HEDFT_CHECK(hipGraphLaunch(c->graph_exec, c->stream));
  }
  for (; done_steps < nsteps; ++done_steps) {
    HEDFT_TRY(steppc_impl(c, deltat));
    HEDFT_TRY(tddft_after_step(c));
  }
  return HEDFT_OK;
}
// non-blocking: nsteps PC iterations queued on the context's stream
int enqueue_pc_async(hedft_ctx* c, double deltat, int nsteps) {
  if (!c || c->mode != HEDFT_MODE_TDDFT || nsteps < 0) return HEDFT_EINVAL;
  if (!c->rk_done) return HEDFT_ESTATE;
  return enqueue_pc(c, deltat, nsteps, 1);
}

// device bytes hedft_create allocates for a cubic grid edge n
size_t ctx_footprint(int n, int mode, int precision) {
  const size_t N = (size_t)n * n * n, Nh = (size_t)(n / 2 + 1) * n * n;
  const size_t fw = (precision == HEDFT_PREC_MIXED) ? 4 : 8;
  const size_t pw = (mode == HEDFT_MODE_TDDFT) ? 16 : 8;
  size_t b = 17 * fw * N + 9 * pw * N + 16 * N + 6 * 2 * fw * Nh + 3 * fw * Nh;
  if (mode == HEDFT_MODE_STATICS) b += 4 * 8 * N;
  else b += 2 * fw * Nh;  // the spectrum whose inverse waits for stage C
  return b + (8u << 20);  // partials, tables, twiddles, impurity block
}
}  // namespace hedft

extern "C" {

int hedft_tddft_run_pc(hedft_ctx* c, double deltat, int32_t nsteps, int32_t use_graph, double* elapsed_ms) {
  if (!c || c->mode != HEDFT_MODE_TDDFT || nsteps < 0) return HEDFT_EINVAL;
  HEDFT_TRY(imp_ready(c));
  if (!c->rk_done) return HEDFT_ESTATE;
  hipSetDevice(c->device);
  if (use_graph && nsteps >= PC_PERIOD) HEDFT_TRY(ensure_graph(c, deltat));  // capture outside the timing
  HEDFT_CHECK(hipEventRecord(c->ev0, c->stream));
  HEDFT_TRY(enqueue_pc(c, deltat, nsteps, use_graph));
  HEDFT_CHECK(hipEventRecord(c->ev1, c->stream));
  HEDFT_CHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HEDFT_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if (elapsed_ms) *elapsed_ms = ms;
  return done(c);
}

int hedft_batch_run_pc(hedft_ctx* const* ctxs, int32_t nctx, const double* deltat, int32_t nsteps,
                       int32_t chunk, double* elapsed_ms) {
  if (!ctxs || nctx <= 0 || !deltat || nsteps < 0 || chunk < 0) return HEDFT_EINVAL;
  const int dev = ctxs[0] ? ctxs[0]->device : -1;
  for (int i = 0; i < nctx; ++i) {
    const hedft_ctx* c = ctxs[i];
    if (!c || c->mode != HEDFT_MODE_TDDFT || c->device != dev) return HEDFT_EINVAL;
    if (!c->rk_done) return HEDFT_ESTATE;
    HEDFT_TRY(imp_ready(c));
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return HEDFT_EINVAL;  // a context appears once
  }
  hipSetDevice(dev);
  const int ch = chunk > 0 ? ((chunk + PC_PERIOD - 1) / PC_PERIOD) * PC_PERIOD : PC_PERIOD;
  for (int i = 0; i < nctx; ++i)
    if (nsteps >= PC_PERIOD) HEDFT_TRY(ensure_graph(ctxs[i], deltat[i]));
  // every context's main stream waits for a common start marker; chunks are
  // issued round-robin so that the droplets advance together, each on its own
  // streams, and their kernels overlap on the device
  hipEvent_t t0 = nullptr, t1 = nullptr;
  HEDFT_CHECK(hipEventCreate(&t0));
  HEDFT_CHECK(hipEventCreate(&t1));
  int rc = HEDFT_OK;
  hipStream_t s0 = ctxs[0]->main_stream;
  if (hipEventRecord(t0, s0) != hipSuccess) rc = HEDFT_EHIP;
  for (int i = 1; i < nctx && rc == HEDFT_OK; ++i)
    if (hipStreamWaitEvent(ctxs[i]->main_stream, t0, 0) != hipSuccess) rc = HEDFT_EHIP;
  for (int done_steps = 0; done_steps < nsteps && rc == HEDFT_OK;) {
    const int k = std::min(ch, nsteps - done_steps);
    for (int i = 0; i < nctx && rc == HEDFT_OK; ++i) rc = enqueue_pc(ctxs[i], deltat[i], k, 1);
    done_steps += k;
  }
  for (int i = 1; i < nctx && rc == HEDFT_OK; ++i) {
    hedft_ctx* c = ctxs[i];
    if (hipEventRecord(c->ev1, c->main_stream) != hipSuccess || hipStreamWaitEvent(s0, c->ev1, 0) != hipSuccess)
      rc = HEDFT_EHIP;
  }
  float ms = 0.f;
  if (rc == HEDFT_OK && (hipEventRecord(t1, s0) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                         hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
    rc = HEDFT_EHIP;
  for (int i = 0; i < nctx; ++i) hipStreamSynchronize(ctxs[i]->main_stream);
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (rc != HEDFT_OK) return rc;
  if (elapsed_ms) *elapsed_ms = ms;
  for (int i = 0; i < nctx; ++i) HEDFT_TRY(done(ctxs[i]));
  return HEDFT_OK;
}

int hedft_tddft_probe_pc(hedft_ctx* c, double deltat, int32_t nsteps, double* avg_ms_pred,
                         double* avg_ms_corr) {
  if (!c || c->mode != HEDFT_MODE_TDDFT || nsteps <= 0 || nsteps > 100000) return HEDFT_EINVAL;
  if (!c->rk_done) return HEDFT_ESTATE;
  hipSetDevice(c->device);
  std::vector<hipEvent_t> ev(4 * (size_t)nsteps, nullptr);
  int rc = HEDFT_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      rc = HEDFT_EHIP;
      break;
    }
  if (rc == HEDFT_OK) {
    c->kt_ev[0] = ev.data();
    c->kt_ev[1] = ev.data() + 2 * (size_t)nsteps;
    c->kt_cap = nsteps;
    c->kt_n[0] = c->kt_n[1] = 0;
    c->kt_on = 1;
    rc = prepare_lap(c);
    if (rc == HEDFT_OK) c->lap_fresh = 1;
    for (int s = 0; s < nsteps && rc == HEDFT_OK; ++s) {
      rc = steppc_impl(c, deltat);
      if (rc == HEDFT_OK) rc = tddft_after_step(c);
    }
    c->kt_on = 0;
    if (rc == HEDFT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = HEDFT_EHIP;
    double* outs[2] = {avg_ms_pred, avg_ms_corr};
    for (int slot = 0; slot < 2 && rc == HEDFT_OK; ++slot) {
      double sum = 0.0;
      for (int k = 0; k < c->kt_n[slot]; ++k) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->kt_ev[slot][2 * k], c->kt_ev[slot][2 * k + 1]) != hipSuccess) {
          rc = HEDFT_EHIP;
          break;
        }
        sum += ms;
      }
      if (outs[slot]) *outs[slot] = c->kt_n[slot] ? sum / c->kt_n[slot] : 0.0;
    }
    c->kt_ev[0] = c->kt_ev[1] = nullptr;
    c->kt_cap = 0;
  }
  for (auto& e : ev)
    if (e) hipEventDestroy(e);
  if (rc != HEDFT_OK) return rc;
  return done(c);
}

static void apply_tuning(hedft_ctx* c, const hedft_tuning* t) {
  c->xy_split_nb[0] = t->xy_split_fwd;
  c->xy_split_nb[1] = t->xy_split_inv;
  c->no_side = t->no_side;
  c->prefer_eager = t->prefer_eager;
  c->lap_late = t->lap_late;
  c->lap25_tz = t->lap25;
  c->fused = t->fused;
  c->lap_fresh = 0;    // the spare psi buffer does not hold the x+y Laplacian
  c->graph_valid = 0;  // the captured graph encodes the launch structure
}

int hedft_set_tuning(hedft_ctx* c, const hedft_tuning* t) {
  if (!c || !t || t->xy_split_fwd < 0 || t->xy_split_inv < 0 || t->lap25 < 0 || t->fused < 0 || t->fused > 1)
    return HEDFT_EINVAL;
  if (t->lap25 > 0 && (c->mode != HEDFT_MODE_TDDFT || lap25_geom(c->d, t->lap25).nblocks == 0)) return HEDFT_EGRID;
  apply_tuning(c, t);
  return HEDFT_OK;
}

int hedft_get_tuning(const hedft_ctx* c, hedft_tuning* t) {
  if (!c || !t) return HEDFT_EINVAL;
  t->xy_split_fwd = c->xy_split_nb[0];
  t->xy_split_inv = c->xy_split_nb[1];
  t->no_side = c->no_side;
  t->prefer_eager = c->prefer_eager;
  t->lap_late = c->lap_late;
  t->lap25 = c->lap25_tz;
  t->fused = c->fused;
  t->ms_per_step = 0.0;
  return HEDFT_OK;
}

int hedft_autotune(hedft_ctx* c, double deltat, int32_t trial_steps, hedft_tuning* best) {
  if (!c || c->mode != HEDFT_MODE_TDDFT || trial_steps < 1) return HEDFT_EINVAL;
  if (!c->rk_done) return HEDFT_ESTATE;
  hipSetDevice(c->device);
  const int k = ((trial_steps + PC_PERIOD - 1) / PC_PERIOD) * PC_PERIOD;
  hedft_tuning cur{}, win{};
  win.ms_per_step = 1e300;
  const int split[4][2] = {{0, 0}, {1, 1}, {1, 0}, {0, 1}};
  for (int sp = 0; sp < 4; ++sp)
    for (int ns = 0; ns < 2; ++ns)
      for (int eg = 0; eg < 2; ++eg)
        for (int ll = 0; ll < 5; ++ll) {
          cur.xy_split_fwd = split[sp][0];
          cur.xy_split_inv = split[sp][1];
          cur.no_side = ns;
          cur.prefer_eager = eg;
          cur.lap_late = ll;
          cur.lap25 = c->lap25_tz;
          cur.fused = c->fused;
          apply_tuning(c, &cur);
          HEDFT_TRY(enqueue_pc(c, deltat, PC_PERIOD, 1));  // warm-up incl. graph capture
          HEDFT_CHECK(hipEventRecord(c->ev0, c->stream));
          HEDFT_TRY(enqueue_pc(c, deltat, k, 1));
          HEDFT_CHECK(hipEventRecord(c->ev1, c->stream));
          HEDFT_CHECK(hipEventSynchronize(c->ev1));
          float ms = 0.f;
          HEDFT_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
          cur.ms_per_step = ms / k;
          if (cur.ms_per_step < win.ms_per_step) win = cur;
        }
  apply_tuning(c, &win);
  if (best) *best = win;
  return done(c);
}

int hedft_last_elapsed(hedft_ctx* c, double* ms) {
  if (!c || !ms) return HEDFT_EINVAL;
  *ms = c->last_ms;
  return HEDFT_OK;
}

}  // extern "C"
