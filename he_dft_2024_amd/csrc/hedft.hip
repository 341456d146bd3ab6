// hedft.hip — the context and the C-ABI around it: creation, fields, uploads
// and downloads, the impurity set-up, the TDDFT driver entry points, graph
// replay of PC iterations, batches and tuning.  The mean field, the
// integrators, the statics steppers and the energies live in meanfield.hip,
// steppers.hip, statics.hip and observables.hip.
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

#include "internal.h"

using namespace hedft;

namespace hedft {

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
int read_scal(hedft_ctx* c, int first, int count, double* out) {
  HEDFT_CHECK(hipMemcpyAsync(c->host_scal, c->scal + first, count * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < count; ++i) out[i] = c->host_scal[i];
  return HEDFT_OK;
}

int done(hedft_ctx* c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "hedft: %s\n", hipGetErrorString(e));
    return HEDFT_EHIP;
  }
  (void)c;
  return HEDFT_OK;
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
  if (c->rem_exec) hipGraphExecDestroy(c->rem_exec);
  if (c->rem_graph) hipGraphDestroy(c->rem_graph);
  void* ptrs[] = {c->den, c->pot4, c->dencg, c->delj4, c->denalf, c->intx, c->inty, c->intz, c->ualphas,
                  c->dxden, c->dyden, c->dzden, c->uext, c->sto1, c->ualph_conv, c->work1, c->work2,
                  c->timec_r, c->hpsi_r, c->psi1s, c->psi2s, c->psil, c->psi, c->psiold[0], c->psiold[1], c->psiold[2], c->hpsiold[0],
                  c->hpsiold[1], c->pc, c->q, c->lap, c->workc1, c->fden, c->spec[0], c->spec[1], c->spec[2],
                  c->spec[3], c->workk1, c->fvlj4, c->wcgk, c->kalfs, c->tw, c->twf, c->lamtab, c->partials, c->scal,
                  c->imp, c->lstop, c->potion, c->potx4, c->specq[0], c->specq[1], c->specd, c->coef2_dev,
                  c->coef1_dev};
  for (void* q : ptrs)
    if (q) hipFree(q);
  for (int k = 0; k < HEDFT_MAX_QIMP; ++k) {
    hedft_ctx::QImp& x = c->qi[k];
    void* qp[] = {x.psix, x.psixold[0], x.psixold[1], x.psixold[2], x.hpsixold[0], x.hpsixold[1], x.pcx, x.qx,
                  x.lapx, x.upotx, x.uimpk};
    for (void* q : qp)
      if (q) hipFree(q);
    for (int m = 0; m < HEDFT_MAX_QIMP; ++m)
      if (c->uimpkk[k][m]) hipFree(c->uimpkk[k][m]);
  }
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
      {
        hedft_ctx::QImp& q = c->qi[c->qsel];
        if (!td) {  // statics: real psix, its potentials and kernel
          switch (f) {
            case HEDFT_PSIX: *ptr = q.psix; *bytes = rN; break;
            case HEDFT_POTX4: *ptr = c->potx4; *bytes = rN; break;
            case HEDFT_UPOTX: *ptr = q.upotx; *bytes = rN; break;
            case HEDFT_DENX: *ptr = c->workc1; *bytes = rN; *computed = 1; break;
            case HEDFT_UIMPK: *ptr = q.uimpk; *bytes = kN; break;
            default: return HEDFT_EINVAL;
          }
          break;
        }
        switch (f) {
          case HEDFT_PSIX: *ptr = q.psix; *bytes = cN; break;
          case HEDFT_POTX4: *ptr = c->potx4; *bytes = rN; break;
          case HEDFT_UPOTX: *ptr = q.upotx; *bytes = rN; break;
          case HEDFT_DENX: *ptr = c->workc1; *bytes = rN; *computed = 1; break;
          case HEDFT_UIMPK: *ptr = q.uimpk; *bytes = kN; break;
          case HEDFT_PCX: *ptr = q.pcx; *bytes = cN; break;
          case HEDFT_QX: *ptr = q.qx; *bytes = cN; break;
          case HEDFT_HPSIXOLD1: case HEDFT_HPSIXOLD2: *ptr = q.hpsixold[f - HEDFT_HPSIXOLD1]; *bytes = cN; break;
          default: *ptr = q.psixold[f - HEDFT_PSIXOLD1]; *bytes = cN; break;
        }
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
      hipLaunchKernelGGL(k_sq<double>, gr, bl, 0, c->stream, c->d.n, rz(c->qi[c->qsel].psix), w);
    else if (field == HEDFT_DENX)
      hipLaunchKernelGGL(k_abs2, gr, bl, 0, c->stream, c->d.n, (const double2*)c->qi[c->qsel].psix, w);
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
  HEDFT_TRY(alpha_term_run(c));
  HEDFT_CHECK(hipMemcpyAsync(c->pot4, c->work2, b, hipMemcpyDeviceToDevice, c->stream));
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
  double s[8], sx[HEDFT_MAX_QIMP] = {0.0};
  HEDFT_TRY(read_scal(c, 0, 8, s));
  if (c->quantum) HEDFT_TRY(read_scal(c, S_ERRPCX, c->quantum, sx));
  // errPsix = sum_k c9 sum|pcx_k| / nxyz (Dynamics_quantum_impurities/steppc.f90:115-139)
  double ex = 0.0;
  for (int k = 0; k < c->quantum; ++k) ex += (9.0 / 121.0) * sx[k];
  if (errHe) *errHe = (9.0 / 121.0) * s[S_ERRPC] / (double)c->d.n;
  if (errimp) *errimp = c->quantum ? ex / (double)c->d.n : c->nimp ? s[S_ERRIMP] : 0.0;
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
  if (c->quantum) HEDFT_TRY(potenimp_quantum(c));
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
  void** reals[] = {&c->qi[0].psix, &c->qi[0].psixold[0], &c->qi[0].psixold[1], (void**)&c->potx4, (void**)&c->qi[0].upotx};
  for (void** f : reals)
    if (!*f) rc = rc ? rc : alloc(c, f, rN);
  if (!c->qi[0].uimpk) rc = rc ? rc : alloc(c, (void**)&c->qi[0].uimpk, kN);
  if (!c->specq[0]) rc = rc ? rc : alloc(c, (void**)&c->specq[0], hN);
  if (rc != HEDFT_OK) return rc;
  HEDFT_CHECK(hipMemcpyAsync(c->qi[0].uimpk, vq, bytes, hipMemcpyHostToDevice, c->stream));
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  c->qi[0].m_imp_u = m_imp_u;
  c->qi[0].h2o2mimp = c->p.h2o2m4 * (4.002602 / m_imp_u);
  c->mu4 = 0.0;
  c->quantum = 1;
  return done(c);
}

// a real half-spectrum kernel into device storage R (FP64 table, converted on the mixed path)
static int upload_kernel(hedft_ctx* c, double* dev, const double* host, size_t bytes) {
  if (c->mixed) {
    HEDFT_CHECK(hipMemcpyAsync(c->workc1, host, bytes, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, c->stream, (long long)c->d.nh,
                       (const double*)c->workc1, (float*)dev);
  } else {
    HEDFT_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
  }
  return HEDFT_OK;
}

// potenimpini (Dynamics_quantum_impurities/potenimpini.f90:19-44): uimp_k(:,:,:,i)
// and uimp_k_k(:,:,:,i,j) of nq impurities, and their masses (steprk.f90:86)
int hedft_set_quantum_impurities(hedft_ctx* c, int32_t nq, const double* uimp_k, const double* uimp_kk,
                                 const int32_t* pair, size_t bytes, const double* m_imp_u, double mu4) {
  if (!c || !uimp_k || !m_imp_u || c->nimp > 0 || nq < 1 || nq > HEDFT_MAX_QIMP) return HEDFT_EINVAL;
  if (bytes != sizeof(double) * (size_t)c->d.nh) return HEDFT_EINVAL;
  for (int k = 0; k < nq; ++k)
    if (!(m_imp_u[k] > 0.0) || (pair && pair[k * nq + k])) return HEDFT_EINVAL;  // uimp_k_k(i,i) is no pair
  if (pair && !uimp_kk) return HEDFT_EINVAL;
  hipSetDevice(c->device);
  HEDFT_CHECK(hipStreamSynchronize(c->stream));  // queued work may still read the old kernels
  if (c->mode == HEDFT_MODE_STATICS) {
    if (nq != 1) return HEDFT_EINVAL;  // He_statics carries one impurity wave function
    c->graph_valid = 0;
    return set_quantum_statics(c, uimp_k, bytes, m_imp_u[0]);
  }
  const size_t fw = c->mixed ? sizeof(float) : sizeof(double);
  const size_t cN = sizeof(double2) * c->d.n, rN = fw * c->d.n, kN = fw * c->d.nh, hN = 2 * fw * c->d.nh;
  const size_t nh = (size_t)c->d.nh;
  int rc = HEDFT_OK;
  if (!c->potx4) rc = rc ? rc : alloc(c, (void**)&c->potx4, rN);
  for (int k = 0; k < 2; ++k)
    if (!c->specq[k]) rc = rc ? rc : alloc(c, (void**)&c->specq[k], hN);
  for (int k = 0; k < nq && rc == HEDFT_OK; ++k) {
    hedft_ctx::QImp& q = c->qi[k];
    void** cplx[] = {&q.psix, &q.psixold[0], &q.psixold[1], &q.psixold[2], &q.hpsixold[0], &q.hpsixold[1],
                     &q.pcx, &q.qx, &q.lapx};
    for (void** f : cplx)
      if (!*f) rc = rc ? rc : alloc(c, f, cN);
    if (!q.upotx) rc = rc ? rc : alloc(c, (void**)&q.upotx, rN);
    if (!q.uimpk) rc = rc ? rc : alloc(c, (void**)&q.uimpk, kN);
    if (rc == HEDFT_OK) rc = upload_kernel(c, q.uimpk, uimp_k + k * nh, bytes);
    for (int m = 0; m < HEDFT_MAX_QIMP && rc == HEDFT_OK; ++m) {
      const bool on = m < nq && m != k && uimp_kk && pair && pair[k + m * nq];
      if (on && !c->uimpkk[k][m]) rc = alloc(c, (void**)&c->uimpkk[k][m], kN);
      if (on && rc == HEDFT_OK) rc = upload_kernel(c, c->uimpkk[k][m], uimp_kk + ((size_t)m * nq + k) * nh, bytes);
      if (!on && c->uimpkk[k][m]) {
        hipFree(c->uimpkk[k][m]);
        c->uimpkk[k][m] = nullptr;
      }
    }
    q.m_imp_u = m_imp_u[k];
    q.h2o2mimp = c->p.h2o2m4 * 4.002603 / m_imp_u[k];
  }
  if (rc != HEDFT_OK) return rc;
  HEDFT_CHECK(hipStreamSynchronize(c->stream));
  c->mu4 = mu4;
  c->quantum = nq;
  c->qsel = 0;
  c->graph_valid = 0;
  c->lap_fresh = 0;
  return done(c);
}

int hedft_set_quantum_impurity(hedft_ctx* c, const double* uimp_k, size_t bytes, double m_imp_u, double mu4) {
  return hedft_set_quantum_impurities(c, 1, uimp_k, nullptr, nullptr, bytes, &m_imp_u, mu4);
}

int hedft_select_quantum_impurity(hedft_ctx* c, int32_t k) {
  if (!c || k < 0 || k >= (c->quantum > 0 ? c->quantum : 1)) return HEDFT_EINVAL;
  c->qsel = k;
  return HEDFT_OK;
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
  for (int k = 0; k < c->quantum; ++k)
    HEDFT_CHECK(hipMemsetAsync(c->qi[k].qx, 0, sizeof(double2) * c->d.n, c->stream));
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

// The host-side state a PC iteration advances: the slot indices, the psi /
// spare pointers the steppers swap, and the side-branch bookkeeping.  A full
// period returns all of it to where it started.
struct HostStep {
  int slots[13];
  void *psi, *lap;
  void *psix[HEDFT_MAX_QIMP], *lapx[HEDFT_MAX_QIMP];
  int side_pre, side_args[9], lap_fresh, ndefer;
  double side_dt;
  size_t side_nblocks;
  double2* defer_spec[2];
  double* defer_out[2];
};
static_assert(sizeof(HostStep) <= sizeof(((hedft_ctx*)nullptr)->rem_pre), "rem_pre holds a HostStep");
static void host_save(const hedft_ctx* c, HostStep& h) {
  std::memset(&h, 0, sizeof h);
  slot_state(c, h.slots);
  h.psi = c->psi;
  h.lap = c->lap;
  for (int k = 0; k < HEDFT_MAX_QIMP; ++k) {
    h.psix[k] = c->qi[k].psix;
    h.lapx[k] = c->qi[k].lapx;
  }
  h.side_pre = c->side_pre;
  std::memcpy(h.side_args, c->side_args, sizeof h.side_args);
  h.lap_fresh = c->lap_fresh;
  h.ndefer = c->ndefer;
  h.side_dt = c->side_dt;
  h.side_nblocks = c->side_nblocks;
  for (int k = 0; k < 2; ++k) {
    h.defer_spec[k] = c->defer_spec[k];
    h.defer_out[k] = c->defer_out[k];
  }
}
static void host_restore(hedft_ctx* c, const HostStep& h) {
  std::memcpy(c->ioldp, h.slots, 3 * sizeof(int));
  std::memcpy(c->ioldh, h.slots + 3, 2 * sizeof(int));
  std::memcpy(c->ioldr, h.slots + 5, 3 * sizeof(int));
  std::memcpy(c->ioldv, h.slots + 8, 3 * sizeof(int));
  std::memcpy(c->iolda, h.slots + 11, 2 * sizeof(int));
  c->psi = h.psi;
  c->lap = h.lap;
  for (int k = 0; k < HEDFT_MAX_QIMP; ++k) {
    c->qi[k].psix = h.psix[k];
    c->qi[k].lapx = h.lapx[k];
  }
  c->side_pre = h.side_pre;
  std::memcpy(c->side_args, h.side_args, sizeof h.side_args);
  c->lap_fresh = h.lap_fresh;
  c->ndefer = h.ndefer;
  c->side_dt = h.side_dt;
  c->side_nblocks = h.side_nblocks;
  for (int k = 0; k < 2; ++k) {
    c->defer_spec[k] = h.defer_spec[k];
    c->defer_out[k] = h.defer_out[k];
  }
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
  // the device-side set-up of the executable graph is done here rather than at
  // its first launch
  HEDFT_CHECK(hipGraphUpload(c->graph_exec, c->stream));
  // capture advanced the host-side slot indices by one full period, i.e. back
  // to where they started
  c->graph_valid = 1;
  c->graph_dt = deltat;
  c->graph_epoch++;
  std::memcpy(c->graph_slots, slots, sizeof slots);
  return HEDFT_OK;
}

// statics: a graph of PC_PERIOD iterations of the driver loop (STEPPCR,
// refresh, poten) from the current slot state, dt and atom number
__global__ void k_copy_scal_mf(double* s, int from, int to) { s[to] = s[from]; }
static int ensure_statics_graph(hedft_ctx* c, double deltat, double n4) {
  int slots[13];
  slot_state(c, slots);
  if (c->graph_valid && c->graph_dt == deltat && c->graph_n4 == n4 &&
      std::memcmp(slots, c->graph_slots, sizeof slots) == 0)
    return HEDFT_OK;
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_valid = 0;
  HEDFT_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = HEDFT_OK;
  for (int s = 0; s < PC_PERIOD && rc == HEDFT_OK; ++s) rc = statics_pc_iter(c, deltat, n4);
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &gr);
  if (rc != HEDFT_OK) {
    if (gr) hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return HEDFT_EHIP;
  c->graph = gr;
  HEDFT_CHECK(hipGraphInstantiate(&c->graph_exec, gr, nullptr, nullptr, 0));
  HEDFT_CHECK(hipGraphUpload(c->graph_exec, c->stream));
  c->graph_valid = 1;
  c->graph_dt = deltat;
  c->graph_n4 = n4;
  c->graph_epoch++;
  std::memcpy(c->graph_slots, slots, sizeof slots);
  return HEDFT_OK;
}

// the remainder graph replays the next nrem iterations from the current state
static bool rem_ready(const hedft_ctx* c, double deltat, int nrem) {
  if (!c->rem_exec || !c->graph_valid || c->rem_epoch != c->graph_epoch || c->rem_dt != deltat ||
      c->rem_n != nrem)
    return false;
  HostStep h;
  host_save(c, h);
  return std::memcmp(&h, c->rem_pre, sizeof h) == 0;
}

// Captures the first nrem (< PC_PERIOD) iterations from the current state
// into the remainder graph, after the period graph (ensure_graph), so that a
// run of q periods + nrem iterations replays graphs only.  The capture
// advances the host-side state by nrem iterations; it is put back, and the
// advanced state is kept for the replay to apply.
static int ensure_rem_graph(hedft_ctx* c, double deltat, int nrem) {
  if (nrem <= 0 || nrem >= PC_PERIOD || !c->graph_valid) return HEDFT_OK;
  if (rem_ready(c, deltat, nrem)) return HEDFT_OK;
  if (c->rem_exec) hipGraphExecDestroy(c->rem_exec);
  if (c->rem_graph) hipGraphDestroy(c->rem_graph);
  c->rem_exec = nullptr;
  c->rem_graph = nullptr;
  HostStep pre, post;
  host_save(c, pre);
  HEDFT_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = HEDFT_OK;
  for (int s = 0; s < nrem && rc == HEDFT_OK; ++s) {
    rc = steppc_impl(c, deltat);
    if (rc == HEDFT_OK) rc = tddft_after_step(c);
  }
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &gr);
  host_save(c, post);
  host_restore(c, pre);
  if (rc != HEDFT_OK) {
    if (gr) hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return HEDFT_EHIP;
  c->rem_graph = gr;
  HEDFT_CHECK(hipGraphInstantiate(&c->rem_exec, gr, nullptr, nullptr, 0));
  HEDFT_CHECK(hipGraphUpload(c->rem_exec, c->stream));
  c->rem_dt = deltat;
  c->rem_n = nrem;
  c->rem_epoch = c->graph_epoch;
  std::memcpy(c->rem_pre, &pre, sizeof pre);
  std::memcpy(c->rem_post, &post, sizeof post);
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
  if (use_graph && done_steps < nsteps && rem_ready(c, deltat, nsteps - done_steps)) {
    HEDFT_CHECK(hipGraphLaunch(c->rem_exec, c->stream));
    HostStep post;
    std::memcpy(&post, c->rem_post, sizeof post);
    host_restore(c, post);
    done_steps = nsteps;
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
  if (use_graph && nsteps >= PC_PERIOD && !c->prefer_eager) {  // captures outside the timing
    HEDFT_TRY(ensure_graph(c, deltat));
    HEDFT_TRY(ensure_rem_graph(c, deltat, nsteps % PC_PERIOD));
  }
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

int hedft_statics_run(hedft_ctx* c, double deltat, int32_t iter0, int32_t niter, double n4, int32_t use_graph,
                      double* mu4, double* mu4err, double* errpc, double* elapsed_ms) {
  if (!c || c->mode != HEDFT_MODE_STATICS || iter0 < 1 || niter < 1 || !(n4 > 0.0)) return HEDFT_EINVAL;
  if (iter0 > 4 && !c->rk_done) return HEDFT_ESTATE;
  hipSetDevice(c->device);
  const int it1 = iter0 + niter;  // one past the last iteration
  const int first_pc = std::max(iter0, 5);
  // PC iterations before the last one replay the graph in whole periods
  const int npc = std::max(0, it1 - first_pc);
  const int q = (use_graph && !c->prefer_eager) ? (npc - 1) / PC_PERIOD : 0;
  if (q > 0 && iter0 > 4) HEDFT_TRY(ensure_statics_graph(c, deltat, n4));  // capture outside the timing
  HEDFT_CHECK(hipEventRecord(c->ev0, c->stream));
  int it = iter0;
  for (; it < it1 && it <= 4; ++it) {
    if (it == it1 - 1) hipLaunchKernelGGL(k_copy_scal_mf, dim3(1), dim3(1), 0, c->stream, c->scal, S_HMEAN, S_MUOLD);
    HEDFT_TRY(statics_rk_iter(c, deltat, n4));
  }
  if (q > 0) {
    HEDFT_TRY(ensure_statics_graph(c, deltat, n4));  // a no-op unless the RK start just ended
    for (int k = 0; k < q; ++k) HEDFT_CHECK(hipGraphLaunch(c->graph_exec, c->stream));
    it += q * PC_PERIOD;
  }
  for (; it < it1; ++it) {
    if (it == it1 - 1) hipLaunchKernelGGL(k_copy_scal_mf, dim3(1), dim3(1), 0, c->stream, c->scal, S_HMEAN, S_MUOLD);
    HEDFT_TRY(statics_pc_iter(c, deltat, n4));
  }
  HEDFT_CHECK(hipEventRecord(c->ev1, c->stream));
  HEDFT_CHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HEDFT_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if (elapsed_ms) *elapsed_ms = ms;
  double sc[5];
  HEDFT_TRY(read_scal(c, 0, 5, sc));
  // mu4err = |1 - mu4/hmean| with mu4 the previous iteration's hmean (steprkr.f90:93, steppcr.f90:104)
  if (mu4err) *mu4err = std::fabs(1.0 - sc[S_MUOLD] / sc[S_HMEAN]);
  if (mu4) *mu4 = sc[S_HMEAN];
  if (errpc) *errpc = (it1 - 1 > 4) ? (9.0 / 121.0) * sc[S_ERRPC] / (double)c->d.n : 0.0;
  return done(c);
}

int hedft_tddft_prepare_pc(hedft_ctx* c, double deltat, int32_t skip, int32_t nsteps) {
  if (!c || c->mode != HEDFT_MODE_TDDFT || skip < 0 || nsteps < 0) return HEDFT_EINVAL;
  HEDFT_TRY(imp_ready(c));
  if (!c->rk_done) return HEDFT_ESTATE;
  if (nsteps < PC_PERIOD || c->prefer_eager) return HEDFT_OK;
  hipSetDevice(c->device);
  HostStep pre;
  host_save(c, pre);
  int rc = HEDFT_OK;
  // the host-side state repeats every PC_PERIOD iterations: bring it to the
  // phase of the run by capturing skip % PC_PERIOD iterations into a graph
  // that is thrown away (capture launches nothing)
  const int sk = skip % PC_PERIOD;
  if (sk > 0) {
    HEDFT_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < sk && rc == HEDFT_OK; ++s) {
      rc = steppc_impl(c, deltat);
      if (rc == HEDFT_OK) rc = tddft_after_step(c);
    }
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    if (gr) hipGraphDestroy(gr);
    if (rc == HEDFT_OK && e != hipSuccess) rc = HEDFT_EHIP;
  }
  if (rc == HEDFT_OK) rc = ensure_graph(c, deltat);
  if (rc == HEDFT_OK) rc = ensure_rem_graph(c, deltat, nsteps % PC_PERIOD);
  host_restore(c, pre);
  return rc;
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
