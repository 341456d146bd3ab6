// common.h — shared device helpers and the context layout of the HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/hedft.h"

#define HEDFT_CHECK(expr)                                                          \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "hedft: HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
              __FILE__, __LINE__);                                                  \
      return HEDFT_EHIP;                                                            \
    }                                                                               \
  } while (0)

#define HEDFT_TRY(expr)      \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != HEDFT_OK) return rc_; \
  } while (0)

namespace hedft {

// ---- complex helpers (interleaved re,im = complex(8)) -------------------------
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by -i (fwd) or +i (inv)
template <bool INV>
__host__ __device__ inline double2 mul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
template <bool INV>
__host__ __device__ inline float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// complex type of a real type (float2 / double2) and a generic constructor
template <typename R>
struct cvec;
template <>
struct cvec<double> {
  using type = double2;
};
template <>
struct cvec<float> {
  using type = float2;
};
template <typename R>
using cplx_t = typename cvec<R>::type;
template <typename C, typename T>
__host__ __device__ inline C mkc(T x, T y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}

// ---- launch geometry -----------------------------------------------------------
struct Dims {
  int nx, ny, nz, nxh;
  long long n, nh;
};

// Stencil coefficients of derivnD for one (id, axis): central row (13) and the
// 12 boundary rows (6 leading + 6 trailing) of the Bickley matrix after the
// boundary-mode rewrite, already divided by h^id.
struct StencilCoef {
  double c[13];
  double rows[12][13];
  int mode;  // 0: boundary rows, 1: periodic, 2: antiperiodic
};

// ---- impurity evaluation on the fly (updatepoten, potenimp.f90:83-123) -------
struct ImpDev {
  int nimp, npot;
  double delinter, rmax;
  const double* potion;  // nimp x npot
  const double* rimp;    // device (nimp,3), k fastest
  int* lstop;            // device flag
  double x0, y0, z0, hx, hy, hz;
};

// Products and sums here are rounded separately (no FMA contraction), as a
// plain CPU build evaluates x(i) = -xmax + h*(i-1) and r; the table lookup is
// discontinuous where r/DelInter is an integer, so r must match to the last bit.
// table node ir and weight rmod of impurity k at (xx, yy, zz)
__device__ inline void imp_node_k(const ImpDev& im, int k, double xx, double yy, double zz, bool flag, int& ir,
                                  double& rmod) {
#pragma clang fp contract(off)
  const double rx = im.rimp[k], ry = im.rimp[k + im.nimp], rz = im.rimp[k + 2 * im.nimp];
  const double zt = (zz - rz) * (zz - rz);
  const double yt = (yy - ry) * (yy - ry) + zt;
  double r = sqrt((xx - rx) * (xx - rx) + yt);
  if (r > im.rmax) {
    if (flag) *im.lstop = 1;
    r = im.rmax;
  }
  // ir = int(r/DelInter), rmod = mod(r,DelInter)/DelInter (potenimp.f90:101-111).
  // mod is the exact remainder: when r/DelInter rounds up to an integer the
  // remainder is just below DelInter while ir already points one node on (the
  // reference behaves so at such points, e.g. grid points exactly 4 A from an
  // impurity sitting on a grid line).
  const double q = trunc(r / im.delinter);
  ir = (int)q;
  if (ir > im.npot - 2) ir = im.npot - 2;
  double rem = fma(-q, im.delinter, r);
  if (rem < 0.0) rem += im.delinter;
  else if (rem >= im.delinter) rem -= im.delinter;
  rmod = rem / im.delinter;
}

// the linear interpolation between the two table values
__device__ inline double imp_lerp(double t0, double t1, double rmod) {
#pragma clang fp contract(off)
  return t0 * (1.0 - rmod) + t1 * rmod;
}

__device__ inline double imp_u_k(const ImpDev& im, int k, double xx, double yy, double zz, bool flag) {
  int ir;
  double rmod;
  imp_node_k(im, k, xx, yy, zz, flag, ir, rmod);
  const double* tab = im.potion + (size_t)k * im.npot;
  return imp_lerp(tab[ir], tab[ir + 1], rmod);
}

// ---- impurity-impurity pair terms (Dynamics_TDDFT/potenimp.f90:40-58,
// energy.f90:168-177) ------------------------------------------------------------
// The reference evaluates Select_pot(drselec_gs_k_k) for the pair force and
// Select_pot(selec_gs_k_k) for the pair energy; here both come as radial
// tables on r_i = i*del (i = 0..np-1), one pair after the other in the order
// (1,2),(1,3),..,(1,N),(2,3),.., read with updatepoten's interpolation
// (potenimp.f90:101-112).  r beyond the last node is clamped to it and
// flagged.
__host__ __device__ inline double radial_lerp(const double* t, int np, double del, double r, int* flag) {
#pragma clang fp contract(off)
  const double rmax = del * (double)(np - 1);
  if (r > rmax) {
    if (flag) *flag = 1;
    r = rmax;
  }
  const double q = trunc(r / del);
  int ir = (int)q;
  if (ir > np - 2) ir = np - 2;
  double rem = fma(-q, del, r);
  if (rem < 0.0) rem += del;
  else if (rem >= del) rem -= del;
  const double rmod = rem / del;
  return t[ir] * (1.0 - rmod) + t[ir + 1] * rmod;
}

// index of pair (k,m), k < m, among the nimp*(nimp-1)/2 pairs
__host__ __device__ inline int pair_index(int k, int m, int nimp) { return k * nimp - k * (k + 1) / 2 + (m - k - 1); }

// ---- impurity block layout on the device (doubles) ---------------------------------
enum {
  IMP_R = 0, IMP_V = 12, IMP_A = 24, IMP_F = 36, IMP_QR = 48, IMP_QV = 60, IMP_PCR = 72,
  IMP_PCV = 84, IMP_ROLD = 96, IMP_VOLD = 132, IMP_AOLD = 168, IMP_MASS = 192, IMP_SIZE = 200
};

// blocks of the grid-stride elementwise kernels (a fixed grid keeps their
// per-block partial sums, and so the reductions, deterministic)
static constexpr int EW_BLOCKS = 1024;

struct PairDev {
  const double* dv;  // force tables (drselec), npairs x np
  int np;
  double del;
  int* lstop;
};

// F_ij(k,m,:) = -D_km(r) (R_k - R_m)/r for m > k, F_ij(m,k,:) = -F_ij(k,m,:)
// (potenimp.f90:40-47); fij[(k*MAXIMP + m)*3 + a]
__host__ __device__ inline void pair_forces(const double* rimp, int nimp, const double* dv, int np, double del,
                                            int* flag, double* fij) {
#pragma clang fp contract(off)
  for (int i = 0; i < HEDFT_MAX_IMP * HEDFT_MAX_IMP * 3; ++i) fij[i] = 0.0;
  for (int k = 0; k < nimp; ++k)
    for (int m = k + 1; m < nimp; ++m) {
      const double dx = rimp[k] - rimp[m], dy = rimp[k + nimp] - rimp[m + nimp],
                   dz = rimp[k + 2 * nimp] - rimp[m + 2 * nimp];
      const double aux1 = sqrt(dx * dx + dy * dy + dz * dz);
      const double s = radial_lerp(dv + (size_t)pair_index(k, m, nimp) * np, np, del, aux1, flag);
      const double d[3] = {dx, dy, dz};
      for (int a = 0; a < 3; ++a) {
        const double f = -s * d[a] / aux1;
        fij[(k * HEDFT_MAX_IMP + m) * 3 + a] = f;
        fij[(m * HEDFT_MAX_IMP + k) * 3 + a] = -f;
      }
    }
}

// grid coordinates x_i = -xmax + h*i, rounded like the CPU evaluation
__device__ inline void imp_xyz(const ImpDev& im, int ix, int iy, int iz, double& xx, double& yy, double& zz) {
#pragma clang fp contract(off)
  xx = im.x0 + im.hx * ix;
  yy = im.y0 + im.hy * iy;
  zz = im.z0 + im.hz * iz;
}

__device__ inline double imp_u(const ImpDev& im, int ix, int iy, int iz) {
  double xx, yy, zz;
  imp_xyz(im, ix, iy, iz, xx, yy, zz);
  double u = 0.0;
  for (int k = 0; k < im.nimp; ++k) u += imp_u_k(im, k, xx, yy, zz, false);
  return u;
}

// ---- LDS fill with memory-level parallelism -------------------------------------
// lds[e] = f(e) for e in [0, total), strided over the NT threads of the block.
// f must load unconditionally from a valid address (callers clamp indices; tile
// slots that no output reads may hold any value): the U loads of a thread are
// then issued back to back before the first LDS store, keeping U requests in
// flight per wavefront instead of one.
template <int U, int NT, typename V, typename F>
__device__ __forceinline__ void lds_fill(V* lds, int total, F f) {
  for (int e0 = threadIdx.x; e0 < total; e0 += NT * U) {
    V tmp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) tmp[u] = f(min(e0 + u * NT, total - 1));
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * NT < total) lds[e0 + u * NT] = tmp[u];
  }
}

// ---- block reduction into per-block partials ----------------------------------
template <int NR>
__device__ inline void block_reduce_store(double (&v)[NR], double* partials) {
  __shared__ double red[NR][32];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double s = v[r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[r][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[threadIdx.x][w];
    const size_t bid = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
    partials[bid * NR + threadIdx.x] = s;
  }
  __syncthreads();
}

}  // namespace hedft

// ---- the context -----------------------------------------------------------------
struct hedft_ctx {
  hedft_grid g;
  hedft_params p;
  int mode;
  int mixed;  // HEDFT_PREC_MIXED: the real mean-field fields, spectra and kernel
              // tables are stored as float (the double* members then point to
              // float storage); psi and its PC history stay double
  int device;
  hipStream_t stream;  // every launch goes to this stream
  hipStream_t main_stream, side_stream;  // stream is main_stream except inside a side branch
  hipEvent_t ev0, ev1, ev_fork, ev_join, ev_mid;
  hedft::Dims d;
  double hx, hy, hz, dxyz, renor;
  // real N fields
  double *den, *pot4, *dencg, *delj4, *denalf, *intx, *inty, *intz, *ualphas;
  double *dxden, *dyden, *dzden, *uext, *sto1, *ualph_conv, *work1, *work2;
  double* timec_r;  // statics local time step
  double* hpsi_r;   // statics H psi
  double *psi1s, *psi2s;  // statics steepest descent: psi one and two steps back
  double* psil;           // statics log-psi steppers: log|psi|
  // psi and PC history: double (statics) or double2 (TDDFT)
  void *psi, *psiold[3], *hpsiold[2], *pc, *q;
  void* lap;        // Laplacian accumulator (same type as psi)
  void* workc1;     // complex N scratch
  // half spectra
  double2 *fden, *spec[4], *workk1;
  // TDDFT: spectra whose inverse waits for stage C (delj4, and upotx with a
  // quantum impurity, read only after the pot4 assembly)
  double2* specd;
  int ndefer;
  double2* defer_spec[2];
  double* defer_out[2];
  double *fvlj4, *wcgk, *kalfs;
  double* tw;       // twiddle tables (complex, interleaved)
  float* twf;       // the same tables rounded to FP32 (mixed path)
  double* lamtab;   // TDDFT absorbing layer: tanh((|x_a|-mean_a)/surf_a) per axis (nx+ny+nz)
  size_t tw_x, tw_xpost, tw_y, tw_z;  // offsets (in complex elements)
  // reductions
  double* partials;
  size_t partials_cap;  // doubles
  double* scal;         // device scalars (see physics.hip)
  double* host_scal;    // pinned mirror
  // stencil coefficient sets: [id-1][axis]
  hedft::StencilCoef coef[2][3];
  // the same with one-sided boundary rows (Icon_local = 0 of the log-psi steppers)
  hedft::StencilCoef coef0[2][3];
  // PC slot bookkeeping (host-side; rkpc: ioldp, ioldh; classicimp: ioldr, ioldv, iolda)
  int ioldp[3], ioldh[2], ioldr[3], ioldv[3], iolda[2];
  int rk_done;
  // lap holds the x+y Laplacian of the current psi (computed on the side
  // stream by the mean-field pass that follows a step); cleared whenever psi
  // may change otherwise
  int lap_fresh;
  // small serial kernels queued by a step for the start of the next mean
  // field's side branch (off the critical path): 1 = impurity predictor,
  // 2 = impurity corrector, 4 = errHe reduction
  int side_pre;
  double side_dt;
  int side_args[9];
  size_t side_nblocks;
  // classical impurities
  int nimp, npot;
  double delinter, rmaxinterpol;
  double* potion;   // device table
  double* imp;      // device impurity state block (see physics.hip layout)
  int* lstop;       // device flag
  double mass[HEDFT_MAX_IMP];
  // impurity-impurity pair tables (hedft_set_impurity_pairs): device force
  // tables, host copy of the energy tables; required when nimp >= 2
  int npair;
  double delpair;
  double* pairtab;      // device: npairs x npair force values (drselec)
  double* pair_v_host;  // host: npairs x npair energy values (selec)
  double last_ms;
  // kernel-duration probe: events around the fused predictor (0) and
  // corrector (1) launches while kt_on
  int kt_on, kt_cap, kt_n[2];
  hipEvent_t* kt_ev[2];  // kt_cap pairs each
  // graph of one PC iteration
  hipGraph_t graph;
  hipGraphExec_t graph_exec;
  double graph_dt;
  int graph_valid;
  int graph_slots[13];  // PC slot indices the graph was captured from
  int graph_epoch;      // incremented by every capture of the period graph
  double graph_n4;      // statics: the atom number the graph was captured with
  // graph of the first rem_n (< period) PC iterations from one slot phase,
  // captured after the period graph of epoch rem_epoch; rem_pre / rem_post:
  // the host-side step state it starts from and leaves (HostStep, hedft.hip)
  hipGraph_t rem_graph;
  hipGraphExec_t rem_exec;
  double rem_dt;
  int rem_n, rem_epoch;
  unsigned char rem_pre[512], rem_post[512];
  // tuning: x+y FFT launches with at most this many arrays use separate row and
  // column passes instead of one whole-plane workgroup per z-plane
  int xy_split_nb[2];  // [forward, inverse]
  int no_side;         // 1: run the side branches on the main stream
  // > 0: the TDDFT steppers compute the x+y Laplacian themselves (k_lap25,
  // z chunks of lap25_tz planes) instead of reading a separate x+y pass
  int lap25_tz;
  hedft::StencilCoef* coef2_dev;  // device copy of coef[1][0..2] (second derivative, x y z)
  hedft::StencilCoef* coef1_dev;  // device copy of coef[0][0..2] (first derivative, x y z)
  // 1: the mean field runs on the main stream alone: the alpha_s divergence
  // with the pot4 assembly in one 2.5-D pass (no x+y pass beside stage C)
  int fused;
  int zconv_fused;     // 1: one workgroup applies all kernels of a convolution
  int prefer_eager;    // 1: hedft_tddft_run_pc / batches skip graph replay
  int lap_late;        // where the next step's x+y Laplacians run (hedft_tuning.lap_late)
  // quantum impurities (Dynamics_quantum_impurities): quantum = N_imp wave
  // functions psix, each with its own PC history, coupled to the helium through
  // its real half-spectrum kernel uimpk (potx4 = sum_k F^-1[F(|psix_k|^2)
  // uimpk_k] enters pot4, upotx_k = F^-1[F(den) uimpk_k] drives psix_k) and to
  // each other through uimpkk (upotx_i += F^-1[F(|psix_j|^2) uimpkk[i][j]]);
  // qsel picks the impurity the PSIX* fields and the centroid refer to
  int quantum, qsel;
  struct QImp {
    double h2o2mimp, m_imp_u;
    void *psix, *psixold[3], *hpsixold[2], *pcx, *qx, *lapx;  // complex N
    double* upotx;                                             // real N (R storage)
    double* uimpk;                                             // real half spectrum (R storage)
  } qi[HEDFT_MAX_QIMP];
  double* uimpkk[HEDFT_MAX_QIMP][HEDFT_MAX_QIMP];  // null: a 'null' pair, or i = j
  double mu4;
  double* potx4;      // real N (R storage)
  double2* specq[2];  // half spectra: |psix|^2, den x uimpk
  // statics (He_statics with limp): psix is real N, psixold[0..1] hold psi1x,
  // psi2x of evolox, specq[0] holds fdenx = F(psix^2) as the driver leaves it
};

namespace hedft {
// field added to pot4 after the correlation term: uext (statics with the
// external-field option) or potx4 (quantum impurity); null when neither
inline double* pot4_ext(const hedft_ctx* c) {
  if (c->quantum) return c->potx4;
  if (c->mode == HEDFT_MODE_STATICS) return c->p.lexternal ? c->uext : nullptr;
  return nullptr;
}

// fft.hip
int fft_setup(hedft_ctx* c);
int fftfw(hedft_ctx* c, const double* in, double2* out);
int fftbk(hedft_ctx* c, double2* in, double* out);

// x-pass prologue (R2C input) and epilogue (C2R output) operations
enum XInOp { XIN_COPY = 0, XIN_CORR = 1, XIN_FGRAD = 2, XIN_DOT3 = 3, XIN_ABS2 = 4 };
enum XOutOp { XOUT_STORE = 0, XOUT_POT4 = 1, XOUT_ACC = 2 };  // ACC: out += value (real-space sums)
// R: storage type of the real fields (double: FP64 path, float: mixed path);
// the pointwise arithmetic of prologues and epilogues is done in double.
template <typename R>
struct XInT {
  int op;
  const R* a[6];
  const double2* z;  // XIN_ABS2: |z|^2 of a complex field
  double c0, c1;
};
template <typename R>
struct XOutT {
  int op;
  R* out;
  const R *delj4, *dencg, *den;
  const R* ext;  // added to pot4 when lext: uext (statics) or potx4 (quantum impurity)
  double a0, a1, csolid, beta, den_m;
  int lsolid, lext;
};
using XIn = XInT<double>;
using XOut = XOutT<double>;
// forward x R2C of up to 4 real inputs into spectra, then y forward
template <typename R>
int fwd_xy(hedft_ctx* c, int nb, const XInT<R>* in, cplx_t<R>* const* spec);
// z pass: optional forward, then for each output o: multiply by kern[o], inverse z.
// nin inputs; kern[i*nout+o], out[i*nout+o].  store_fwd[i] (may be null)
// receives the full forward spectrum.
template <typename R>
int zconv(hedft_ctx* c, int nin, cplx_t<R>* const* in, int nout, const R* const* kern,
          cplx_t<R>* const* out, int do_fwd, cplx_t<R>* const* store_fwd);
// y inverse then x C2R with epilogue for nb spectra
template <typename R>
int inv_yx(hedft_ctx* c, int nb, cplx_t<R>* const* spec, const XOutT<R>* out);

// hedft.hip: queue nsteps PC iterations without waiting (graph replays for
// whole periods); device bytes of one context
int enqueue_pc_async(hedft_ctx* c, double deltat, int nsteps);
size_t ctx_footprint(int n, int mode, int precision);

// impurity.hip: force reduction F = -sum(partials) dxyz + pair forces [, aimp = F/m]
ImpDev make_impdev(hedft_ctx* c);
double* force_partials(hedft_ctx* c);
int force_finalize(hedft_ctx* c, size_t nblocks, int update_aimp, int nr);
int npairs_of(int nimp);
void clear_pairs(hedft_ctx* c);
// several classical impurities need their pair tables: HEDFT_ESTATE until given
int imp_ready(const hedft_ctx* c);

// stencil.hip
void stencil_setup(hedft_ctx* c);
enum SInOp { SIN_PLAIN = 0, SIN_FINT = 1 };
// plain derivnD of a field (real or complex) along axis (0..2)
int deriv_plain(hedft_ctx* c, int id, int axis, const void* in, void* out, int is_complex);
// grad(den) -> dxden,dyden,dzden with impurity force partials (forceimp)
// grad(den) and, with want_force, the impurity force partials; finalize = 0
// leaves their reduction to derden_force_finalize (queued later on the stream)
int derden_force(hedft_ctx* c, int want_force, int update_aimp, int finalize = 1);
int derden_force_finalize(hedft_ctx* c, int update_aimp);
// divergence of (1-denalf/den0s)*int_a, then ualphas and pot4 assembly
int alpha_div_xy(hedft_ctx* c);      // x+y divergence of f int_a into work1
int alpha_div_z_pot4(hedft_ctx* c);  // z divergence + ualphas + pot4 assembly
// the whole divergence (x, y, z) + ualphas + pot4 assembly in one 2.5-D pass;
// results equal those of alpha_div_xy followed by alpha_div_z_pot4
int alpha_div25_pot4(hedft_ctx* c);
// 1 when the grid takes the 2.5-D tiles (lap25_geom)
int div25_ok(const hedft_ctx* c);
// x+y Laplacian of psi into out (default c->lap); the z pass is the fused step kernel
int lap_xy(hedft_ctx* c, const void* psi, int is_complex, void* out = nullptr);
}  // namespace hedft
