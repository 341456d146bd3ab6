// impurity.hip — forceimp of the classical impurities: the reduction of the
// per-block force partials, the impurity-impurity pair force, and the
// standalone forceimp entry point (Dynamics_TDDFT/potenimp.f90:23-67).
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "common.h"

namespace hedft {

// F_k,a = -sum_blocks(partials[3k+a]) * dxyz + sum_m F_ij(k,m,a); optionally
// aimp = F/m (DFT4He3d.f90:692-696).  Per-block partials hold NR = 3 values per
// accumulated impurity.  Each thread first sums its blocks into NR independent
// registers (loads in flight together), then one tree over the workgroup.
template <int NR>
__global__ void __launch_bounds__(256) k_force_finalize(const double* __restrict__ part, int nblocks,
                                                        const double* rimp, double* fout, double* aout,
                                                        const double* mass, int nimp, double dxyz,
                                                        int update_aimp, PairDev pr) {
  __shared__ double sh[NR][256];
  double s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] += part[(size_t)b * NR + r];
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] = s[r];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // F(k,:) = -sum(grad(den) uimp_k) dxyz + sum_m F_ij(k,m,:)  (potenimp.f90:49-58)
    double fij[HEDFT_MAX_IMP * HEDFT_MAX_IMP * 3];
    const bool pairs = nimp > 1 && pr.dv;
    if (pairs) pair_forces(rimp, nimp, pr.dv, pr.np, pr.del, pr.lstop, fij);
    for (int k = 0; k < nimp && 3 * k < NR; ++k)
      for (int a = 0; a < 3; ++a) {
        double f = -sh[3 * k + a][0] * dxyz;
        if (pairs)
          for (int m = 0; m < nimp; ++m) f += fij[(k * HEDFT_MAX_IMP + m) * 3 + a];
        fout[k + nimp * a] = f;
        if (update_aimp) aout[k + nimp * a] = f / mass[k];
      }
  }
}

// forceimp's He part as an elementwise pass over a stored gradient
// (potenimp.f90:49-52): per-block sums of d_a(den) uimp_k
template <typename R>
__global__ void __launch_bounds__(256) k_force_ew(Dims d, ImpDev im, const R* gx, const R* gy, const R* gz,
                                                  double* partials) {
  double red[3 * HEDFT_MAX_IMP];
  for (int r = 0; r < 3 * HEDFT_MAX_IMP; ++r) red[r] = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < d.n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % d.nx), iy = (int)((i / d.nx) % d.ny), iz = (int)(i / ((long long)d.nx * d.ny));
    double xx, yy, zz;
    imp_xyz(im, ix, iy, iz, xx, yy, zz);
    for (int k = 0; k < im.nimp; ++k) {
      const double u = imp_u_k(im, k, xx, yy, zz, false);
      red[3 * k] += (double)gx[i] * u;
      red[3 * k + 1] += (double)gy[i] * u;
      red[3 * k + 2] += (double)gz[i] * u;
    }
  }
  block_reduce_store<3 * HEDFT_MAX_IMP>(red, partials);
}

ImpDev make_impdev(hedft_ctx* c) {
  ImpDev im;
  im.nimp = c->nimp;
  im.npot = c->npot;
  im.delinter = c->delinter;
  im.rmax = c->rmaxinterpol;
  im.potion = c->potion;
  im.rimp = c->imp + IMP_R;
  im.lstop = c->lstop;
  im.x0 = -c->g.xmax;
  im.y0 = -c->g.ymax;
  im.z0 = -c->g.zmax;
  im.hx = c->hx;
  im.hy = c->hy;
  im.hz = c->hz;
  return im;
}

double* force_partials(hedft_ctx* c) { return c->partials + c->partials_cap / 8; }

int npairs_of(int nimp) { return nimp * (nimp - 1) / 2; }

void clear_pairs(hedft_ctx* c) {
  if (c->pairtab) hipFree(c->pairtab);
  std::free(c->pair_v_host);
  c->pairtab = nullptr;
  c->pair_v_host = nullptr;
  c->npair = 0;
}

int imp_ready(const hedft_ctx* c) { return (c->nimp >= 2 && !c->pairtab) ? HEDFT_ESTATE : HEDFT_OK; }

static PairDev make_pairdev(hedft_ctx* c) {
  PairDev p;
  p.dv = c->pairtab;
  p.np = c->npair;
  p.del = c->delpair;
  p.lstop = c->lstop;
  return p;
}

int force_finalize(hedft_ctx* c, size_t nblocks, int update_aimp, int nr) {
  double* imp = c->imp;
  if (nr == 3)
    hipLaunchKernelGGL(k_force_finalize<3>, dim3(1), dim3(256), 0, c->stream, force_partials(c), (int)nblocks,
                       imp + IMP_R, imp + IMP_F, imp + IMP_A, imp + IMP_MASS, c->nimp, c->dxyz, update_aimp,
                       make_pairdev(c));
  else
    hipLaunchKernelGGL(k_force_finalize<3 * HEDFT_MAX_IMP>, dim3(1), dim3(256), 0, c->stream, force_partials(c),
                       (int)nblocks, imp + IMP_R, imp + IMP_F, imp + IMP_A, imp + IMP_MASS, c->nimp, c->dxyz,
                       update_aimp, make_pairdev(c));
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

// ---- scratch of the allocation-free forceimp ----------------------------------
// One copy per GPU, alive until the process ends: the per-block force partials
// of k_force_ew (EW_BLOCKS blocks x 3 HEDFT_MAX_IMP sums) and one flag word that
// takes a clamped pair distance when the caller passes no flag of its own.
// Every call on a GPU shares that GPU's copy, so calls that overlap on one GPU
// race.  Each block's partials are written in full before the finalize pass
// reads them, so nothing needs clearing between calls.
static constexpr size_t NOALLOC_PARTIALS = (size_t)EW_BLOCKS * 3 * HEDFT_MAX_IMP;
static constexpr size_t NOALLOC_DOUBLES = NOALLOC_PARTIALS + 32;
static constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

static std::mutex scratch_mu;
static double* scratch_dev[NOALLOC_MAX_DEVICES];  // device address of `scratch` per GPU

// The array's name is a host-side stand-in: its device address comes from
// hipGetSymbolAddress, once per GPU with that GPU current.
static int scratch_on(int device, double** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return HEDFT_EINVAL;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "hedft: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return HEDFT_EHIP;
    }
    scratch_dev[device] = static_cast<double*>(p);
  }
  *out = scratch_dev[device];
  return HEDFT_OK;
}

}  // namespace hedft

using namespace hedft;

extern "C" {

int hedft_set_impurity_pairs(hedft_ctx* c, int32_t npair, double delpair, const double* vpair,
                             const double* dvpair) {
  if (!c || c->nimp < 2 || npair < 2 || !(delpair > 0.0) || !vpair || !dvpair) return HEDFT_EINVAL;
  hipSetDevice(c->device);
  HEDFT_CHECK(hipStreamSynchronize(c->stream));  // queued work may still read the old tables
  c->graph_valid = 0;
  clear_pairs(c);
  const size_t cnt = (size_t)npairs_of(c->nimp) * npair;
  c->pair_v_host = (double*)std::malloc(sizeof(double) * cnt);
  if (!c->pair_v_host) return HEDFT_ENOMEM;
  if (hipMalloc(&c->pairtab, sizeof(double) * cnt) != hipSuccess) {
    c->pairtab = nullptr;
    clear_pairs(c);
    return HEDFT_ENOMEM;
  }
  std::memcpy(c->pair_v_host, vpair, sizeof(double) * cnt);
  HEDFT_CHECK(hipMemcpy(c->pairtab, dvpair, sizeof(double) * cnt, hipMemcpyHostToDevice));
  c->npair = npair;
  c->delpair = delpair;
  return HEDFT_OK;
}

// standalone forceimp() from the stored gradient (potenimp.f90:23-67)
int hedft_forceimp(hedft_ctx* c) {
  if (!c) return HEDFT_EINVAL;
  if (c->nimp == 0) return HEDFT_OK;
  HEDFT_TRY(imp_ready(c));
  if (c->mixed)
    hipLaunchKernelGGL(k_force_ew<float>, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, make_impdev(c),
                       reinterpret_cast<const float*>(c->dxden), reinterpret_cast<const float*>(c->dyden),
                       reinterpret_cast<const float*>(c->dzden), force_partials(c));
  else
    hipLaunchKernelGGL(k_force_ew<double>, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d, make_impdev(c),
                       c->dxden, c->dyden, c->dzden, force_partials(c));
  return force_finalize(c, EW_BLOCKS, 0, 3 * HEDFT_MAX_IMP);
}

int hedft_forceimp_noalloc(const hedft_grid* g, int32_t device, void* stream, const double* dxden,
                           const double* dyden, const double* dzden, int32_t nimp, int32_t npot, double delinter,
                           double rmaxinterpol, const double* potion, const double* rimp, int32_t npair,
                           double delpair, const double* dvpair, double* force, int32_t* lstop) {
  if (!g || g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || !dxden || !dyden || !dzden || !rimp || !force)
    return HEDFT_EINVAL;
  if (nimp < 1 || nimp > HEDFT_MAX_IMP || npot < 2 || !potion || !(delinter > 0.0)) return HEDFT_EINVAL;
  if (nimp >= 2 && (npair < 2 || !(delpair > 0.0) || !dvpair)) return HEDFT_EINVAL;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return HEDFT_EHIP;
  if (hipSetDevice(device) != hipSuccess) return HEDFT_EINVAL;
  double* sc = nullptr;
  int rc = scratch_on(device, &sc);
  if (rc == HEDFT_OK) {
    const hipStream_t st = static_cast<hipStream_t>(stream);
    Dims d;
    d.nx = g->nx;
    d.ny = g->ny;
    d.nz = g->nz;
    d.nxh = g->nx / 2 + 1;
    d.n = (long long)d.nx * d.ny * d.nz;
    d.nh = (long long)d.nxh * d.ny * d.nz;
    const double hx = 2.0 * std::fabs(g->xmax) / d.nx, hy = 2.0 * std::fabs(g->ymax) / d.ny,
                 hz = 2.0 * std::fabs(g->zmax) / d.nz;
    int* flag = lstop ? lstop : reinterpret_cast<int*>(sc + NOALLOC_PARTIALS);
    ImpDev im;
    im.nimp = nimp;
    im.npot = npot;
    im.delinter = delinter;
    im.rmax = rmaxinterpol;
    im.potion = potion;
    im.rimp = rimp;
    im.lstop = flag;
    im.x0 = -g->xmax;
    im.y0 = -g->ymax;
    im.z0 = -g->zmax;
    im.hx = hx;
    im.hy = hy;
    im.hz = hz;
    PairDev pr;
    pr.dv = nimp >= 2 ? dvpair : nullptr;
    pr.np = npair;
    pr.del = delpair;
    pr.lstop = flag;
    hipLaunchKernelGGL(k_force_ew<double>, dim3(EW_BLOCKS), dim3(256), 0, st, d, im, dxden, dyden, dzden, sc);
    hipLaunchKernelGGL(k_force_finalize<3 * HEDFT_MAX_IMP>, dim3(1), dim3(256), 0, st, sc, EW_BLOCKS, rimp, force,
                       nullptr, nullptr, nimp, hx * hy * hz, 0, pr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      fprintf(stderr, "hedft: %s\n", hipGetErrorString(e));
      rc = HEDFT_EHIP;
    }
  }
  hipSetDevice(prev);
  return rc;
}

}  // extern "C"
