// internal.h — what the translation units of libhedft share: the device
// scalar slots, the integrator constants, the per-block partial-sum
// reduction, the absorbing-layer operand of H psi, field-pointer helpers for
// the storage type, and the functions one unit calls in another.
#pragma once
#include <vector>

#include "stencil_kernel.h"

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
  S_ERRPCX = 32,   // sum|pcx_k| of quantum impurity k (HEDFT_MAX_QIMP slots)
  S_ENERGY = 8,  // 8 partial energy sums
  S_ENERGYQ = 16,  // quantum-impurity energy sums: 2 per impurity, then one per pair (14)
  S_OBS = 16,      // 13 observable sums (used alone)
  NSCAL = 40
};

static const double RK_A[4] = {0.5, 1.0 - 0.70710678118654752440, 1.0 + 0.70710678118654752440, 1.0 / 6.0};
static const double RK_B[4] = {2.0, 1.0, 1.0, 2.0};
static const double RK_C[4] = {0.5, 1.0 - 0.70710678118654752440, 1.0 + 0.70710678118654752440, 0.5};
static const double C112 = 112.0 / 121.0, C9 = 9.0 / 121.0, C4O3 = 4.0 / 3.0;

static inline size_t partials_region(hedft_ctx* c) { return c->partials_cap / 8; }
static inline double* gen_partials(hedft_ctx* c) { return c->partials; }
// psix corrector of quantum impurity k: regions 2 .. 2 + HEDFT_MAX_QIMP - 1 of the 8
static inline double* x_partials(hedft_ctx* c, int k = 0) { return c->partials + (2 + k) * partials_region(c); }

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
static XOutT<R> xo_acc(double* out) {
  XOutT<R> o{};
  o.op = XOUT_ACC;
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

static inline double2* cz(void* p) { return (double2*)p; }
static inline double* rz(void* p) { return (double*)p; }

// out = a^2 elementwise (the statics impurity density psix^2)
template <typename T>
__global__ void k_sq(long long n, const T* a, T* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = a[i] * a[i];
}

// hedft.hip
int done(hedft_ctx* c);  // reports (and clears) a launch error
int read_scal(hedft_ctx* c, int first, int count, double* out);  // device scalars, after a stream sync
// meanfield.hip
void kernel_tables(const hedft_ctx* c, std::vector<double>& fv, std::vector<double>& wc, std::vector<double>& ka);
int poten(hedft_ctx* c, int want_force);
int poten_tddft(hedft_ctx* c, int want_force, int with_lap = 0);
int poten_statics(hedft_ctx* c);
int statics_refresh(hedft_ctx* c);
int statics_rk_iter(hedft_ctx* c, double dt, double n4);
int statics_pc_iter(hedft_ctx* c, double dt, double n4);
int alpha_term_run(hedft_ctx* c);    // term_alfa on its own (stages B and C of alpha_s)
int potenimp_quantum(hedft_ctx* c);  // upotx and potx4 of a quantum impurity
// F^-1[F(|psix_k|^2) kern] into out (added to out with acc != 0)
int qconv_run(hedft_ctx* c, double* out, int k, const double* kern, int acc);
// steppers.hip
void flush_side_errors(hedft_ctx* c);
void flush_side_impurity(hedft_ctx* c);
void flush_side_pre(hedft_ctx* c);
Lam make_lam(hedft_ctx* c);
int prepare_lap(hedft_ctx* c);
int steprk_impl(hedft_ctx* c, double dt);
int steppc_impl(hedft_ctx* c, double dt);
int tddft_after_step(hedft_ctx* c);

}  // namespace hedft
