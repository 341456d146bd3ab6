// observables.hip — energies (energy.f90 of both programs) and the TDDFT
// driver's observables and quantum-impurity centroid.
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
// sum |grad psix_k|^2, and (with_conv) sum conv den with conv = sum_k
// F^-1[F(|psix_k|^2) uimp_k]
template <typename R>
struct OutEnergyX {
  static constexpr int NR = 2;
  double* partials;
  const double* acc;
  const R *conv, *den;
  int with_conv;
  __device__ __forceinline__ void store(size_t i, int, int, int, double2 v, double* red) const {
    red[0] += acc[i] + vsq<double2>(v);
    if (with_conv) red[1] += (double)conv[i] * (double)den[i];
  }
};

// per-block sums of a[i] |z_i|^2 over a fixed grid
template <typename R>
__global__ void __launch_bounds__(256) k_dot_abs2(long long n, const R* a, const double2* z, double* partials) {
  double red[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double2 v = z[i];
    red[0] += (double)a[i] * (v.x * v.x + v.y * v.y);
  }
  block_reduce_store<1>(red, partials);
}

// sk[k] = sum |grad psix_k|^2, *sh = sum conv den, eij[p] = sum
// F^-1[F(|psix_k|^2) uimp_k_k(k,m)] |psix_m|^2 for the p-th pair k < m with a
// kernel (energy.f90:102-129; the X-He potential summed in real space term by
// term instead of in the spectrum, equal to rounding); returns the pair count
template <typename R>
static int energy_sums_x(hedft_ctx* c, double* sk, double* sh, double* eij, int* npairs) {
  const int nq = c->quantum;
  for (int k = 0; k < nq; ++k) HEDFT_TRY(qconv_run(c, c->work1, k, c->qi[k].uimpk, k > 0));
  for (int k = 0; k < nq; ++k) {
    const double2* px = (const double2*)c->qi[k].psix;
    launch_pencil_xy<double2, true>(c->stream, c->d, c->coef[0][0], c->coef[0][1], InPlain<double2>{px},
                                    OutSqXY<double2>{nullptr, (double*)c->workc1});
    OutEnergyX<R> o;
    o.partials = gen_partials(c);
    o.acc = (const double*)c->workc1;
    o.conv = fld<R>(c->work1);
    o.den = fld<R>(c->den);
    o.with_conv = (k == 0);
    launch_pencil<double2, 2>(c->stream, c->d, c->coef[0][2], InPlain<double2>{px}, o);
    hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(256), 0, c->stream, gen_partials(c),
                       (int)pencil_geom<double2>(c->d, 2).nblocks, c->scal + S_ENERGYQ + 2 * k, 1.0);
  }
  int np = 0;
  for (int k = 0; k < nq; ++k)
    for (int m = k + 1; m < nq; ++m) {
      if (!c->uimpkk[k][m]) continue;
      HEDFT_TRY(qconv_run(c, c->work2, k, c->uimpkk[k][m], 0));
      hipLaunchKernelGGL(k_dot_abs2<R>, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, fld<R>(c->work2),
                         (const double2*)c->qi[m].psix, gen_partials(c));
      finalize<1>(c, gen_partials(c), EW_BLOCKS, S_ENERGYQ + 2 * HEDFT_MAX_QIMP + np, 1.0);
      ++np;
    }
  double s[2 * HEDFT_MAX_QIMP + HEDFT_MAX_QIMP * (HEDFT_MAX_QIMP - 1) / 2];
  HEDFT_TRY(read_scal(c, S_ENERGYQ, 2 * HEDFT_MAX_QIMP + np, s));
  for (int k = 0; k < nq; ++k) sk[k] = s[2 * k];
  *sh = s[1];
  for (int p = 0; p < np; ++p) eij[p] = s[2 * HEDFT_MAX_QIMP + p];
  *npairs = np;
  return HEDFT_OK;
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
  const double* px = rz(c->qi[0].psix);
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

extern "C" {

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
      double sk[HEDFT_MAX_QIMP], sh, eij[HEDFT_MAX_QIMP * (HEDFT_MAX_QIMP - 1) / 2];
      int np = 0;
      if (c->mixed) HEDFT_TRY(energy_sums_x<float>(c, sk, &sh, eij, &np));
      else HEDFT_TRY(energy_sums_x<double>(c, sk, &sh, eij, &np));
      double ek = 0.0, ei = 0.0;
      for (int k = 0; k < c->quantum; ++k) ek += c->qi[k].h2o2mimp * sk[k];
      for (int q = 0; q < np; ++q) ei += eij[q];
      e->ekinx = ek * dv;
      e->eHeX = sh * dv;
      e->eimpu_impu = ei * dv;
    }
    e->eimpu = e->ekinx + e->eHeX + e->eimpu_impu;
  } else if (c->quantum) {
    // He_statics/energy.f90:76-93,137-149: etot = etot4 + ekinx + eimpu
    double sk;
    HEDFT_TRY(energy_sums_xs(c, &sk));
    e->ekinx = c->qi[0].h2o2mimp * sk * dv;
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

int hedft_quantum_centroid(hedft_ctx* c, double* norm, double* rcm) {
  if (!c || !c->quantum) return HEDFT_EINVAL;
  hipLaunchKernelGGL(k_centroid, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, (const double2*)c->qi[c->qsel].psix,
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

}  // extern "C"
