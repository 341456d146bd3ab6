// fft.hip — 3-D R2C / C2R FFTs for gfx950 as LDS-staged Stockham passes.
//
// Replaces the FFTW plans of fft.f90 (He_statics/fft.f90:5-89,
// Dynamics_TDDFT/fft.f90:5-168): forward R2C unnormalised with e^{-i}; C2R
// scaled by renor = 1/(nx ny nz) after the inverse.  Arrays keep the
// reference layout: real (nx,ny,nz) x fastest, half spectrum (nx/2+1,ny,nz).
//
// Passes:
//   x  : rows of nx reals packed as nx/2 complex, Stockham FFT of length nx/2
//        in LDS, then the real-to-half-complex split (forward) or its inverse
//        before the FFT (backward).  Prologue / epilogue ops fuse the pointwise
//        work that feeds or consumes a transform (correlation source, alpha_s
//        products, pot4 assembly).
//   y  : tiles of TX consecutive kx columns x ny, FFT along y in LDS.
//   z  : same along z; the convolution variant runs forward z, multiplies by
//        one to three spectral kernels and runs inverse z while the column
//        tile stays in LDS (the spectrum of a convolution never round-trips
//        HBM between the forward and inverse z transforms).
#include <cmath>
#include <cstdlib>

#include "common.h"

namespace hedft {

// ---------------------------------------------------------------------------
// radix-R DFT in registers on C = float2 / double2; INV selects e^{+i}
// ---------------------------------------------------------------------------
template <bool INV, typename C>
__device__ __forceinline__ void dft4(C& u0, C& u1, C& u2, C& u3) {
  C t0 = cadd(u0, u2), t1 = csub(u0, u2), t2 = cadd(u1, u3), t3 = mul_mi<INV>(csub(u1, u3));
  u0 = cadd(t0, t2);
  u2 = csub(t0, t2);
  u1 = cadd(t1, t3);
  u3 = csub(t1, t3);
}

template <bool INV, typename C>
__device__ __forceinline__ void dft8(C* u) {
  using T = decltype(u[0].x);
  C e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
  C o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  const T r = (T)0.70710678118654752440;
  if (!INV) {
    o1 = mkc<C>((o1.x + o1.y) * r, (o1.y - o1.x) * r);
    o3 = mkc<C>((o3.y - o3.x) * r, (-o3.x - o3.y) * r);
  } else {
    o1 = mkc<C>((o1.x - o1.y) * r, (o1.x + o1.y) * r);
    o3 = mkc<C>((-o3.x - o3.y) * r, (o3.x - o3.y) * r);
  }
  o2 = mul_mi<INV>(o2);
  u[0] = cadd(e0, o0); u[4] = csub(e0, o0);
  u[1] = cadd(e1, o1); u[5] = csub(e1, o1);
  u[2] = cadd(e2, o2); u[6] = csub(e2, o2);
  u[3] = cadd(e3, o3); u[7] = csub(e3, o3);
}

template <bool INV, typename C>
__device__ __forceinline__ void dft3(C* u) {
  using T = decltype(u[0].x);
  const T s3 = (T)0.86602540378443864676;
  C s = cadd(u[1], u[2]);
  C t = mkc<C>(u[0].x - (T)0.5 * s.x, u[0].y - (T)0.5 * s.y);
  C d = mul_mi<INV>(csub(u[1], u[2]));
  d = cscale(d, s3);
  u[0] = cadd(u[0], s);
  u[1] = cadd(t, d);
  u[2] = csub(t, d);
}

template <bool INV, typename C>
__device__ __forceinline__ void dft5(C* u) {
  using T = decltype(u[0].x);
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  C a1 = cadd(u[1], u[4]), d1 = csub(u[1], u[4]);
  C a2 = cadd(u[2], u[3]), d2 = csub(u[2], u[3]);
  C y0 = cadd(u[0], cadd(a1, a2));
  C p1 = mkc<C>(u[0].x + c1 * a1.x + c2 * a2.x, u[0].y + c1 * a1.y + c2 * a2.y);
  C p2 = mkc<C>(u[0].x + c2 * a1.x + c1 * a2.x, u[0].y + c2 * a1.y + c1 * a2.y);
  C b1 = mul_mi<INV>(mkc<C>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
  C b2 = mul_mi<INV>(mkc<C>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
  u[0] = y0;
  u[1] = cadd(p1, b1);
  u[4] = csub(p1, b1);
  u[2] = cadd(p2, b2);
  u[3] = csub(p2, b2);
}

// 16 = 4 x 4: u[4 n1 + n2] -> X[k1 + 4 k2]
template <bool INV, typename C>
__device__ __forceinline__ void dft16(C* u) {
  using T = decltype(u[0].x);
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, r = (T)0.70710678118654752440;
  C a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    a[n2][0] = u[n2];
    a[n2][1] = u[4 + n2];
    a[n2][2] = u[8 + n2];
    a[n2][3] = u[12 + n2];
    dft4<INV>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  // twiddles W16^(n2 k1) (conjugated for the inverse)
  const T sg = INV ? (T)1 : (T)-1;
  auto tw = [&](C v, T wr, T wi) { return mkc<C>(v.x * wr - v.y * wi, v.x * wi + v.y * wr); };
  a[1][1] = tw(a[1][1], c1, sg * s1);
  a[1][2] = tw(a[1][2], r, sg * r);
  a[1][3] = tw(a[1][3], s1, sg * c1);
  a[2][1] = tw(a[2][1], r, sg * r);
  a[2][2] = mul_mi<INV>(a[2][2]);
  a[2][3] = tw(a[2][3], -r, sg * r);
  a[3][1] = tw(a[3][1], s1, sg * c1);
  a[3][2] = tw(a[3][2], -r, sg * r);
  a[3][3] = tw(a[3][3], -c1, -sg * s1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    C b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
    dft4<INV>(b0, b1, b2, b3);
    u[k1] = b0;
    u[k1 + 4] = b1;
    u[k1 + 8] = b2;
    u[k1 + 12] = b3;
  }
}

template <int R, bool INV, typename C>
__device__ __forceinline__ void dft(C* u) {
  if constexpr (R == 2) {
    C a = u[0];
    u[0] = cadd(a, u[1]);
    u[1] = csub(a, u[1]);
  } else if constexpr (R == 3) {
    dft3<INV>(u);
  } else if constexpr (R == 4) {
    dft4<INV>(u[0], u[1], u[2], u[3]);
  } else if constexpr (R == 5) {
    dft5<INV>(u);
  } else if constexpr (R == 16) {
    dft16<INV>(u);
  } else {
    static_assert(R == 8, "radix");
    dft8<INV>(u);
  }
}

// ---------------------------------------------------------------------------
// One Stockham autosort pass (radix R, P = product of earlier radices) over a
// column of length L held in LDS with element stride `stride`, executed by T
// threads (t = 0..T-1).  tw[m] = exp(-2 pi i m / L).
// ---------------------------------------------------------------------------
template <int L, int R, int P, bool INV, int T, typename C>
__device__ __forceinline__ void spass(C* col, int stride, int t, const C* __restrict__ tw) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
  C u[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int i = t + b * T;
    if (PER * T == NB || i < NB) {
#pragma unroll
      for (int j = 0; j < R; ++j) u[b][j] = col[(i + j * NB) * stride];
      if constexpr (P > 1) {
        const int k = i % P;
#pragma unroll
        for (int j = 1; j < R; ++j) {
          C w = tw[(k * j) * (L / (P * R))];
          if (INV) w.y = -w.y;
          u[b][j] = cmul(u[b][j], w);
        }
      }
      dft<R, INV>(u[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int i = t + b * T;
    if (PER * T == NB || i < NB) {
      const int k = i % P;
      const int base = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) col[(base + j * P) * stride] = u[b][j];
    }
  }
  __syncthreads();
}

template <int L, bool INV, int T, int P, int R, int... Rest, typename C>
__device__ __forceinline__ void run_passes(C* col, int stride, int t, const C* tw) {
  spass<L, R, P, INV, T>(col, stride, t, tw);
  if constexpr (sizeof...(Rest) > 0) run_passes<L, INV, T, P * R, Rest...>(col, stride, t, tw);
}

// The last pass of a column FFT with its results handed to put(position, value)
// (a global store) instead of written back to LDS.
template <int L, int R, int P, bool INV, int T, typename C, typename Put>
__device__ __forceinline__ void spass_out(const C* col, int stride, int t, const C* __restrict__ tw, Put put) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int i = t + b * T;
    if (PER * T == NB || i < NB) {
      C u[R];
#pragma unroll
      for (int j = 0; j < R; ++j) u[j] = col[(i + j * NB) * stride];
      if constexpr (P > 1) {
        const int k = i % P;
#pragma unroll
        for (int j = 1; j < R; ++j) {
          C w = tw[(k * j) * (L / (P * R))];
          if (INV) w.y = -w.y;
          u[j] = cmul(u[j], w);
        }
      }
      dft<R, INV>(u);
      const int k = i % P;
      const int base = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) put(base + j * P, u[j]);
    }
  }
}

template <int L, bool INV, int T, int P, int R, int... Rest, typename C, typename Put>
__device__ __forceinline__ void run_passes_to(C* col, int stride, int t, const C* tw, Put put) {
  if constexpr (sizeof...(Rest) == 0) {
    spass_out<L, R, P, INV, T>(col, stride, t, tw, put);
  } else {
    spass<L, R, P, INV, T>(col, stride, t, tw);
    run_passes_to<L, INV, T, P * R, Rest...>(col, stride, t, tw, put);
  }
}

// First pass (radix R1 = L / T, no twiddles) on operands already in registers:
// thread t holds elements t + j T, j < R1, and writes its results to the
// Stockham positions t R1 + j.
template <int L, bool INV, int T, int R1, typename C>
__device__ __forceinline__ void first_pass_regs(C (&u)[R1], C* col, int stride, int t) {
  static_assert(L / R1 == T, "one first-pass butterfly per thread");
  dft<R1, INV>(u);
#pragma unroll
  for (int j = 0; j < R1; ++j) col[(t * R1 + j) * stride] = u[j];
  __syncthreads();
}

template <int L, bool INV, int T, int R1, int... Rest, typename C>
__device__ __forceinline__ void run_regs_lds(C (&u)[R1], C* col, int stride, int t, const C* tw) {
  first_pass_regs<L, INV, T>(u, col, stride, t);
  run_passes<L, INV, T, R1, Rest...>(col, stride, t, tw);
}

template <int L, bool INV, int T, int R1, int... Rest, typename C, typename Put>
__device__ __forceinline__ void run_regs_out(C (&u)[R1], C* col, int stride, int t, const C* tw, Put put) {
  first_pass_regs<L, INV, T>(u, col, stride, t);
  run_passes_to<L, INV, T, R1, Rest...>(col, stride, t, tw, put);
}

// Stockham pass over C columns of length L at once, all NT threads of the
// block sharing the C*L/R butterflies.  Element e of column c lives at
// base[c*cs + e*es].  CF: consecutive threads take consecutive columns
// (columns interleaved in LDS) instead of consecutive butterflies of one column
// (columns contiguous), so that a wavefront reads consecutive LDS addresses.
template <int L, int R, int P, bool INV, int NT, int NC, bool CF, typename C>
__device__ __forceinline__ void mpass(C* base, int cs, int es, const C* __restrict__ tw) {
  constexpr int NB = L / R, ITEMS = NC * NB, PER = (ITEMS + NT - 1) / NT;
  C u[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int item = threadIdx.x + b * NT;
    if (PER * NT == ITEMS || item < ITEMS) {
      const int c = CF ? item % NC : item / NB, i = CF ? item / NC : item % NB;
      const C* col = base + c * cs;
#pragma unroll
      for (int j = 0; j < R; ++j) u[b][j] = col[(i + j * NB) * es];
      if constexpr (P > 1) {
        const int k = i % P;
#pragma unroll
        for (int j = 1; j < R; ++j) {
          C w = tw[(k * j) * (L / (P * R))];
          if (INV) w.y = -w.y;
          u[b][j] = cmul(u[b][j], w);
        }
      }
      dft<R, INV>(u[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int item = threadIdx.x + b * NT;
    if (PER * NT == ITEMS || item < ITEMS) {
      const int c = CF ? item % NC : item / NB, i = CF ? item / NC : item % NB;
      C* col = base + c * cs;
      const int k = i % P;
      const int o = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) col[(o + j * P) * es] = u[b][j];
    }
  }
  __syncthreads();
}

template <int L, bool INV, int NT, int NC, bool CF, int P, int R, int... Rest, typename C>
__device__ __forceinline__ void run_mpasses(C* base, int cs, int es, const C* tw) {
  mpass<L, R, P, INV, NT, NC, CF>(base, cs, es, tw);
  if constexpr (sizeof...(Rest) > 0) run_mpasses<L, INV, NT, NC, CF, P * R, Rest...>(base, cs, es, tw);
}

// Factorisations (first radix largest keeps the thread count T = L / 8).
template <int L>
struct Rad;
#define HEDFT_RAD(LEN, TT, ...)                                                          \
  template <>                                                                            \
  struct Rad<LEN> {                                                                      \
    static constexpr int T = TT;                                                         \
    template <bool INV, typename C>                                                      \
    __device__ static __forceinline__ void run(C* c, int s, int t, const C* tw) {       \
      run_passes<LEN, INV, TT, 1, __VA_ARGS__>(c, s, t, tw);                             \
    }                                                                                    \
    template <bool INV, int NT, int NC, bool CF, typename C>                             \
    __device__ static __forceinline__ void run_multi(C* b, int cs, int es, const C* tw) { \
      run_mpasses<LEN, INV, NT, NC, CF, 1, __VA_ARGS__>(b, cs, es, tw);                  \
    }                                                                                    \
    /* first pass from registers u[j] = element t + j T, result in LDS */              \
    template <bool INV, typename C>                                                      \
    __device__ static __forceinline__ void run_from_regs(C (&u)[LEN / TT], C* c, int s, int t, \
                                                         const C* tw) {                  \
      run_regs_lds<LEN, INV, TT, __VA_ARGS__>(u, c, s, t, tw);                          \
    }                                                                                    \
    /* first pass from registers, last pass to put(position, value) */                 \
    template <bool INV, typename C, typename Put>                                        \
    __device__ static __forceinline__ void run_regs_to(C (&u)[LEN / TT], C* c, int s, int t, \
                                                       const C* tw, Put put) {           \
      run_regs_out<LEN, INV, TT, __VA_ARGS__>(u, c, s, t, tw, put);                     \
    }                                                                                    \
  };
HEDFT_RAD(16, 4, 4, 4)
HEDFT_RAD(24, 3, 8, 3)
HEDFT_RAD(32, 4, 8, 4)
HEDFT_RAD(48, 6, 8, 3, 2)
HEDFT_RAD(64, 8, 8, 8)
HEDFT_RAD(96, 12, 8, 4, 3)
HEDFT_RAD(128, 16, 8, 4, 4)
HEDFT_RAD(192, 24, 8, 8, 3)
HEDFT_RAD(256, 32, 8, 8, 4)
#undef HEDFT_RAD

// Factorisations of the column FFTs inside the whole-plane kernels: fewer,
// wider passes (each pass is a full LDS round trip of the plane and a barrier).
template <int L>
struct RadP : Rad<L> {};
template <>
struct RadP<128> {
  template <bool INV, int NT, int NC, bool CF, typename C>
  __device__ static __forceinline__ void run_multi(C* b, int cs, int es, const C* tw) {
    run_mpasses<128, INV, NT, NC, CF, 1, 16, 8>(b, cs, es, tw);
  }
};

// ---------------------------------------------------------------------------
// x-pass prologue / epilogue (arithmetic in double, storage in R)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ double xin_eval(const XInT<R>& in, size_t i) {
  switch (in.op) {
    case XIN_CORR: {  // den*dencg*(cp4+cpp4*dencg)   (poten.f90:166-168)
      const double cg = in.a[1][i];
      return (double)in.a[0][i] * cg * (in.c0 + in.c1 * cg);
    }
    case XIN_FGRAD:  // falfs*d_a(den), falfs = 1 - denalf/den0s  (term_alfa.f90:26-34)
      return (1.0 - (double)in.a[0][i] / in.c0) * (double)in.a[1][i];
    case XIN_DOT3:  // grad(den).int_alf  (term_alfa.f90:61-65)
      return (double)in.a[0][i] * in.a[3][i] + (double)in.a[1][i] * in.a[4][i] +
             (double)in.a[2][i] * in.a[5][i];
    case XIN_ABS2: {  // psix*conjg(psix)  (Dynamics_quantum_impurities/potenimp.f90:23)
      const double2 v = in.z[i];
      return v.x * v.x + v.y * v.y;
    }
    default:
      return in.a[0][i];
  }
}

template <typename R>
__device__ __forceinline__ void xout_apply(const XOutT<R>& o, size_t i, double v) {
  if (o.op == XOUT_POT4) {
    // pot4 = delj4 + dencg^2 (cp4/2 + cpp4/3 dencg) + conv [+ uext | potx4] [+ penalty]
    // (He_statics/poten.f90:73-91,111-115; Dynamics_TDDFT/poten.f90:194-218;
    // Dynamics_quantum_impurities/poten.f90:116-121)
    const double cg = o.dencg[i];
    double p = (double)o.delj4[i] + cg * cg * (o.a0 + o.a1 * cg) + v;
    if (o.lext) p += o.ext[i];
    if (o.lsolid) {
      const double dn = o.den[i];
      const double th = tanh(o.beta * (dn - o.den_m));
      p += o.csolid * (1.0 + th + o.beta * dn * (1.0 - th * th));
    }
    o.out[i] = (R)p;
  } else if (o.op == XOUT_ACC) {
    o.out[i] = (R)((double)o.out[i] + v);
  } else {
    o.out[i] = (R)v;
  }
}

constexpr int MAXB = 4;  // arrays per batched launch
template <typename R>
struct XInBatch {
  XInT<R> in[MAXB];
  cplx_t<R>* out[MAXB];
};
template <typename R>
struct XOutBatch {
  const cplx_t<R>* in[MAXB];
  XOutT<R> out[MAXB];
};
template <typename R>
struct SpecBatch {
  cplx_t<R>* p[MAXB];
};
template <typename R>
struct ZConvArgs {
  const cplx_t<R>* in[MAXB];
  cplx_t<R>* out[MAXB * 3];
  cplx_t<R>* store[MAXB];
  const R* kern[MAXB * 3];  // kern[i*NOUT + o]
  int do_fwd;
};

// Calls use(getv) with the loader of the x-pass prologue operation, getv(i)
// returning the reals i, i+1 of the transformed field as one complex (the
// operation is uniform per workgroup, so each gets one branch-free load
// sequence).
template <typename R, typename Use>
__device__ __forceinline__ void with_xin_pairs(const XInT<R>& in, Use use) {
  using C = cplx_t<R>;
  auto ld2 = [](const R* p, size_t i) { return *reinterpret_cast<const C*>(p + i); };
  switch (in.op) {
    case XIN_CORR:
      use([&](size_t i) {
        const C dn = ld2(in.a[0], i), cg = ld2(in.a[1], i);
        return mkc<C>((R)((double)dn.x * cg.x * (in.c0 + in.c1 * cg.x)),
                      (R)((double)dn.y * cg.y * (in.c0 + in.c1 * cg.y)));
      });
      break;
    case XIN_FGRAD:
      use([&](size_t i) {
        const C al = ld2(in.a[0], i), gr = ld2(in.a[1], i);
        return mkc<C>((R)((1.0 - (double)al.x / in.c0) * gr.x), (R)((1.0 - (double)al.y / in.c0) * gr.y));
      });
      break;
    case XIN_DOT3:
      use([&](size_t i) {
        const C g0 = ld2(in.a[0], i), g1 = ld2(in.a[1], i), g2 = ld2(in.a[2], i);
        const C i0 = ld2(in.a[3], i), i1 = ld2(in.a[4], i), i2 = ld2(in.a[5], i);
        return mkc<C>((R)((double)g0.x * i0.x + (double)g1.x * i1.x + (double)g2.x * i2.x),
                      (R)((double)g0.y * i0.y + (double)g1.y * i1.y + (double)g2.y * i2.y));
      });
      break;
    case XIN_ABS2:
      use([&](size_t i) {
        const double2 z0 = in.z[i], z1 = in.z[i + 1];
        return mkc<C>((R)(z0.x * z0.x + z0.y * z0.y), (R)(z1.x * z1.x + z1.y * z1.y));
      });
      break;
    default:
      use([&](size_t i) { return ld2(in.a[0], i); });
  }
}

// ---------------------------------------------------------------------------
// fused x + y passes of one z-plane held entirely in LDS (NXH x NY complex,
// up to 160 KB on gfx950): forward = rows R2C then columns FFT; inverse =
// columns inverse FFT then rows C2R.  One workgroup per (plane, array).
// ---------------------------------------------------------------------------
// X_k of the half-complex split: ((z + zc) - i W^k (z - zc)) / 2 with zc = conj(Z_{M-k})
template <typename C>
__device__ __forceinline__ C split_one(C z, C zc, C w) {
  using T = decltype(z.x);
  const C ev = cscale(cadd(z, zc), (T)0.5);
  const C wo = cmul(w, cscale(csub(z, zc), (T)0.5));
  return mkc<C>(ev.x + wo.y, ev.y - wo.x);
}

// Z'_k of the c2r merge: (x + conj(xm)) + i conj(W^k) (x - conj(xm))
template <typename C>
__device__ __forceinline__ C merge_one(C x, C xm, C w) {
  const C xmc = cconj(xm);
  const C sm = cadd(x, xmc);
  const C wd = cmul(cconj(w), csub(x, xmc));
  return mkc<C>(sm.x - wd.y, sm.y + wd.x);
}

template <int NX, int NY, typename R>
struct Plane {
  static constexpr int M = NX / 2, NXH = M + 1, PAD = NXH;  // odd row pitch
  static constexpr int ITX = NY * (M / 2), ITY = NXH * (NY / 2);
  static constexpr int NT0 = (ITX > ITY ? ITX : ITY) / 4;
  static constexpr int NT = NT0 >= 1024 ? 1024 : (NT0 <= 64 ? 64 : ((NT0 + 63) / 64) * 64);
  static constexpr size_t LDS = sizeof(cplx_t<R>) * NY * PAD;
};

template <int NX, int NY, typename R>
__global__ void __launch_bounds__((Plane<NX, NY, R>::NT)) __attribute__((amdgpu_waves_per_eu(1, 4)))
    k_xyfwd(XInBatch<R> a, Dims d, const cplx_t<R>* __restrict__ twx, const cplx_t<R>* __restrict__ twp,
            const cplx_t<R>* __restrict__ twy) {
  using C = cplx_t<R>;
  using PL = Plane<NX, NY, R>;
  constexpr int M = PL::M, NXH = PL::NXH, PAD = PL::PAD, NT = PL::NT;
  extern __shared__ unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);
  const XInT<R>& in = a.in[blockIdx.y];
  C* out = a.out[blockIdx.y];
  const size_t row0 = (size_t)blockIdx.x * NY;
  auto fill = [&](auto getv) {
    for (int e0 = threadIdx.x; e0 < NY * M; e0 += NT * 4) {
      C v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + u * NT, NY * M - 1);
        const int y = e / M, j = e - y * M;
        v[u] = getv((row0 + y) * NX + 2 * j);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + u * NT;
        if (e < NY * M) {
          const int y = e / M, j = e - y * M;
          P[y * PAD + j] = v[u];
        }
      }
    }
  };
  with_xin_pairs(in, fill);
  __syncthreads();
  Rad<M>::template run_multi<false, NT, NY, true>(P, PAD, 1, twx);
  // real-to-half-complex split of every row, in place: X_k, k = 0..M, from
  // Z_k = x_2k + i x_2k+1 (Z_M = Z_0).  One thread owns the pair (k, M-k), so
  // the split needs no register staging and no extra barrier.
  constexpr int MH = M / 2 + 1;
  for (int e = threadIdx.x; e < NY * MH; e += NT) {
    const int y = e / MH, k = e - y * MH;
    C* row = P + y * PAD;
    const C a0 = row[k], b0 = row[k == 0 ? 0 : M - k];
    const C xk = split_one(a0, cconj(b0), twp[k]);
    const C xm = split_one(b0, cconj(a0), twp[M - k]);
    row[k] = xk;
    row[M - k] = xm;
  }
  __syncthreads();
  RadP<NY>::template run_multi<false, NT, NXH, true>(P, 1, PAD, twy);
  C* o = out + (size_t)blockIdx.x * NY * NXH;
  for (int e = threadIdx.x; e < NY * NXH; e += NT) {
    const int y = e / NXH, k = e - y * NXH;
    o[e] = P[y * PAD + k];
  }
}

template <int NX, int NY, typename R>
__global__ void __launch_bounds__((Plane<NX, NY, R>::NT)) __attribute__((amdgpu_waves_per_eu(1, 4)))
    k_xyinv(XOutBatch<R> a, Dims d, const cplx_t<R>* __restrict__ twx, const cplx_t<R>* __restrict__ twp,
            const cplx_t<R>* __restrict__ twy, double renor) {
  using C = cplx_t<R>;
  using PL = Plane<NX, NY, R>;
  constexpr int M = PL::M, NXH = PL::NXH, PAD = PL::PAD, NT = PL::NT;
  extern __shared__ unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);
  const C* in = a.in[blockIdx.y] + (size_t)blockIdx.x * NY * NXH;
  const XOutT<R>& out = a.out[blockIdx.y];
  static_assert(PAD == NXH, "plane rows are stored contiguously");
  lds_fill<8, NT>(P, NY * NXH, [&](int e) { return in[e]; });
  __syncthreads();
  RadP<NY>::template run_multi<true, NT, NXH, true>(P, 1, PAD, twy);
  // half-complex-to-real merge, in place: Z'_k = (X_k + X*_{M-k}) + i W^-k
  // (X_k - X*_{M-k}), k = 0..M-1, imaginary parts of X_0 and X_M dropped (c2r
  // semantics).  One thread owns the pair (k, M-k).
  constexpr int MH = M / 2 + 1;
  for (int e = threadIdx.x; e < NY * MH; e += NT) {
    const int y = e / MH, k = e - y * MH;
    C* row = P + y * PAD;
    C a0 = row[k], b0 = row[M - k];
    if (k == 0) {
      a0.y = 0;
      b0.y = 0;
    }
    const C zk = merge_one(a0, b0, twp[k]);
    const C zm = merge_one(b0, a0, twp[M - k]);
    row[k] = zk;
    if (k != 0) row[M - k] = zm;
  }
  __syncthreads();
  Rad<M>::template run_multi<true, NT, NY, true>(P, PAD, 1, twx);
  const size_t row0 = (size_t)blockIdx.x * NY;
  for (int e = threadIdx.x; e < NY * M; e += NT) {
    const int y = e / M, j = e - y * M;
    const C v = P[y * PAD + j];
    const size_t i = (row0 + y) * NX + 2 * j;
    xout_apply(out, i, (double)v.x * renor);
    xout_apply(out, i + 1, (double)v.y * renor);
  }
}

// ---------------------------------------------------------------------------
// 128 x 128 planes with fewer LDS round trips.  The x passes read their
// operands from global memory straight into registers (forward) or write
// their results straight from registers (inverse), and the y pass next to
// global memory does the same, so the plane crosses LDS 6-7 times per
// transform instead of 12.  Lanes 8c..8c+7 own row c: the 8 x 8 exchange
// between the two radix-8 x passes, the real-pair split and the c2r merge
// only read the lane's own row, which lies in its own wavefront, so they are
// ordered by wave-scope fences instead of workgroup barriers.  Between the
// two x passes the Stockham position 8a + b of a row is stored at 8a + (b ^ a)
// (an XOR swizzle that spreads both the radix-8 writes and reads over the
// LDS banks).  Arithmetic per element is that of k_xyfwd / k_xyinv.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(1, 4)))
    k_xyfwd128(XInBatch<R> a, const cplx_t<R>* __restrict__ twx, const cplx_t<R>* __restrict__ twp,
               const cplx_t<R>* __restrict__ twy) {
  using C = cplx_t<R>;
  constexpr int NX = 128, NY = 128, M = 64, NXH = 65, PAD = 65, NT = 1024;
  extern __shared__ unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);
  const XInT<R>& in = a.in[blockIdx.y];
  const int c = threadIdx.x >> 3, i = threadIdx.x & 7;
  // reals 2(i + 8j), 2(i + 8j) + 1 of row c sit at g0 + 16 j
  const size_t g0 = ((size_t)blockIdx.x * NY + c) * NX + 2 * i;
  C u[8];
  auto load = [&](auto getv) {
#pragma unroll
    for (int j = 0; j < 8; ++j) u[j] = getv(g0 + 16 * j);
  };
  with_xin_pairs(in, load);
  // x pass 1 (radix 8, no twiddles): result j of butterfly i is Stockham position 8i + j
  dft<8, false>(u);
  C* row = P + c * PAD;
#pragma unroll
  for (int j = 0; j < 8; ++j) row[8 * i + (j ^ i)] = u[j];
  wave_sync();
  // x pass 2 (radix 8): butterfly i takes positions i + 8j; its results are Z_{i+8j}
#pragma unroll
  for (int j = 0; j < 8; ++j) u[j] = row[8 * j + (i ^ j)];
#pragma unroll
  for (int j = 1; j < 8; ++j) u[j] = cmul(u[j], twx[i * j]);
  dft<8, false>(u);
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; ++j) row[i + 8 * j] = u[j];
  wave_sync();
  // real-to-half-complex split X_k = split(Z_k, conj Z_{M-k}, W^k) for k = i + 8j
  // (Z_M = Z_0); lane i = 0 also forms X_M from Z_0
  C x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = i + 8 * j;
    x[j] = split_one(u[j], cconj(row[(M - k) & (M - 1)]), twp[k]);
  }
  const C xnyq = split_one(u[0], cconj(u[0]), twp[M]);
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; ++j) row[i + 8 * j] = x[j];
  if (i == 0) row[M] = xnyq;
  __syncthreads();
  // y pass 1 (radix 16) within LDS
  mpass<NY, 16, 1, false, NT, NXH, true>(P, 1, PAD, twy);
  // y pass 2 (radix 8, after 16): from LDS to the half spectrum
  constexpr int RR = 8, PP = 16, NB = NY / RR, ITEMS = NXH * NB, PER = (ITEMS + NT - 1) / NT;
  C* o = a.out[blockIdx.y] + (size_t)blockIdx.x * NY * NXH;
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int item = threadIdx.x + b * NT;
    if (item < ITEMS) {
      const int cc = item % NXH, ii = item / NXH;
      C v[RR];
#pragma unroll
      for (int j = 0; j < RR; ++j) v[j] = P[cc + (ii + j * NB) * PAD];
      const int k = ii % PP;
#pragma unroll
      for (int j = 1; j < RR; ++j) v[j] = cmul(v[j], twy[(k * j) * (NY / (PP * RR))]);
      dft<RR, false>(v);
      const int oo = (ii - k) * RR + k;
#pragma unroll
      for (int j = 0; j < RR; ++j) o[(size_t)(oo + j * PP) * NXH + cc] = v[j];
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(1, 4)))
    k_xyinv128(XOutBatch<R> a, const cplx_t<R>* __restrict__ twx, const cplx_t<R>* __restrict__ twp,
               const cplx_t<R>* __restrict__ twy, double renor) {
  using C = cplx_t<R>;
  constexpr int NX = 128, NY = 128, M = 64, NXH = 65, PAD = 65, NT = 1024;
  extern __shared__ unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);
  const C* in = a.in[blockIdx.y] + (size_t)blockIdx.x * NY * NXH;
  const XOutT<R>& out = a.out[blockIdx.y];
  // y pass 1 (radix 16, no twiddles) from the half spectrum: butterfly ii of
  // column cc takes rows ii + 8j; result j is Stockham position 16 ii + j
  {
    constexpr int NB = NY / 16, ITEMS = NXH * NB;
    const int item = threadIdx.x;
    if (item < ITEMS) {
      const int cc = item % NXH, ii = item / NXH;
      C v[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = in[(ii + j * NB) * NXH + cc];
      dft<16, true>(v);
#pragma unroll
      for (int j = 0; j < 16; ++j) P[cc + (16 * ii + j) * PAD] = v[j];
    }
  }
  __syncthreads();
  // y pass 2 (radix 8, after 16) within LDS
  mpass<NY, 8, 16, true, NT, NXH, true>(P, 1, PAD, twy);
  // c2r merge fused into x pass 1: Z'_k = merge(X_k, X_{M-k}, W^k), k = i + 8j,
  // imaginary parts of X_0 and X_M dropped
  const int c = threadIdx.x >> 3, i = threadIdx.x & 7;
  C* row = P + c * PAD;
  C u[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = i + 8 * j;
    C a0 = row[k], b0 = row[M - k];
    if (k == 0) {
      a0.y = 0;
      b0.y = 0;
    }
    u[j] = merge_one(a0, b0, twp[k]);
  }
  dft<8, true>(u);
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; ++j) row[8 * i + (j ^ i)] = u[j];
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; ++j) u[j] = row[8 * j + (i ^ j)];
#pragma unroll
  for (int j = 1; j < 8; ++j) {
    C w = twx[i * j];
    w.y = -w.y;
    u[j] = cmul(u[j], w);
  }
  dft<8, true>(u);
  // z_{i+8j} = (x_{2(i+8j)}, x_{2(i+8j)+1}) of row c, through the epilogue
  const size_t g0 = ((size_t)blockIdx.x * NY + c) * NX + 2 * i;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    xout_apply(out, g0 + 16 * j, (double)u[j].x * renor);
    xout_apply(out, g0 + 16 * j + 1, (double)u[j].y * renor);
  }
}

// ---------------------------------------------------------------------------
// x R2C: rows of nx = 2M reals -> M+1 complex
// ---------------------------------------------------------------------------
template <int M, int ROWS, typename R>
__global__ void __launch_bounds__(Rad<M>::T * ROWS)
    k_xr2c(XInBatch<R> a, Dims d, const cplx_t<R>* __restrict__ tw, const cplx_t<R>* __restrict__ twp) {
  using C = cplx_t<R>;
  constexpr int T = Rad<M>::T, R1 = M / T;
  __shared__ C lds[ROWS][M + 1];
  const XInT<R>& in = a.in[blockIdx.y];
  C* out = a.out[blockIdx.y];
  const int nrows = d.ny * d.nz;
  const int t = threadIdx.x % T, r = threadIdx.x / T;
  const int row = blockIdx.x * ROWS + r;
  // the first pass's operands (pairs t + j T of the row) straight from global
  // memory; rows past the end load zeros and store nothing
  C u[R1];
#pragma unroll
  for (int j = 0; j < R1; ++j) {
    u[j] = mkc<C>((R)0, (R)0);
    if (row < nrows) {
      const size_t i = (size_t)row * (2 * M) + 2 * (t + j * T);
      if (in.op == XIN_COPY) {
        u[j] = *reinterpret_cast<const C*>(in.a[0] + i);
      } else {
        u[j].x = (R)xin_eval(in, i);
        u[j].y = (R)xin_eval(in, i + 1);
      }
    }
  }
  Rad<M>::template run_from_regs<false>(u, &lds[r][0], 1, t, tw);
  if (row < nrows) {
    for (int k = t; k <= M; k += T) {
      const C zk = lds[r][k == M ? 0 : k];
      const C zm = cconj(lds[r][k == 0 ? 0 : M - k]);
      out[(size_t)row * (M + 1) + k] = split_one(zk, zm, twp[k]);
    }
  }
}

// ---------------------------------------------------------------------------
// x C2R: M+1 complex -> nx = 2M reals (imaginary parts of X_0, X_M ignored).
// The c2r merge of the pair (k, M-k) is formed from global memory straight
// into the first pass's registers, and the last pass writes the reals through
// the epilogue.
// ---------------------------------------------------------------------------
template <int M, int ROWS, typename R>
__global__ void __launch_bounds__(Rad<M>::T * ROWS)
    k_xc2r(XOutBatch<R> a, Dims d, const cplx_t<R>* __restrict__ tw, const cplx_t<R>* __restrict__ twp,
           double renor) {
  using C = cplx_t<R>;
  constexpr int T = Rad<M>::T, R1 = M / T;
  __shared__ C lds[ROWS][M + 1];
  const C* in = a.in[blockIdx.y];
  const XOutT<R>& out = a.out[blockIdx.y];
  const int nrows = d.ny * d.nz;
  const int t = threadIdx.x % T, r = threadIdx.x / T;
  const int row = blockIdx.x * ROWS + r;
  const bool live = row < nrows;
  const C* h = in + (size_t)min(row, nrows - 1) * (M + 1);
  C u[R1];
#pragma unroll
  for (int j = 0; j < R1; ++j) {
    const int k = t + j * T;
    C xk = h[k];
    C xm = h[M - k];
    if (k == 0) {
      xk.y = 0;
      xm.y = 0;
    }
    u[j] = merge_one(xk, xm, twp[k]);
  }
  const size_t o = (size_t)min(row, nrows - 1) * (2 * M);
  Rad<M>::template run_regs_to<true>(u, &lds[r][0], 1, t, tw, [&](int z, C v) {
    if (live) {
      xout_apply(out, o + 2 * z, (double)v.x * renor);
      xout_apply(out, o + 2 * z + 1, (double)v.y * renor);
    }
  });
}

// ---------------------------------------------------------------------------
// column passes over a half spectrum.  The columns of one axis are numbered
// flat: y pass c = kx + NXH*kz (a tile of TX consecutive c is contiguous
// except across a plane boundary), z pass c = kx + NXH*ky (contiguous).  TX
// divides nothing in particular: the last tile is partial only at the very
// end, so no tile is wasted on the single Nyquist column of every row.
// ---------------------------------------------------------------------------
// 16 columns per tile at every length (256 B of contiguous complex double
// per row; at L = 256 a 512-thread tile of 64 KB measured 10 % faster on the
// 256^3 mixed step than 8 columns)
template <int L>
struct ColTile {
  static constexpr int T = Rad<L>::T;
  static constexpr int TX = 16;
  static constexpr int NT = T * TX;
};

template <int AX>
__device__ __forceinline__ size_t col_addr(const Dims& d, long long c, size_t& stride) {
  if (AX == 1) {
    stride = d.nxh;
    const long long z = c / d.nxh, kx = c - z * d.nxh;
    return (size_t)z * d.nxh * d.ny + kx;
  }
  stride = (size_t)d.nxh * d.ny;
  return (size_t)c;
}

template <int L, int AX, bool INV, typename R>
__global__ void __launch_bounds__(ColTile<L>::NT)
    k_colfft(SpecBatch<R> a, Dims d, long long ncol, const cplx_t<R>* __restrict__ tw) {
  using C = cplx_t<R>;
  constexpr int TX = ColTile<L>::TX, T = Rad<L>::T, R1 = L / T;
  __shared__ C lds[L * TX];
  C* s = a.p[blockIdx.y];
  // thread (column xl0, t0) loads the first pass's operands t0 + j T straight
  // from global memory and the last pass stores straight back (in place: the
  // workgroup owns its columns and has read them all before the first barrier)
  const int xl0 = threadIdx.x % TX, t0 = threadIdx.x / TX;
  const long long c0 = (long long)blockIdx.x * TX;
  const long long col = min(c0 + xl0, ncol - 1);  // clamped columns are never stored
  const bool live = c0 + xl0 < ncol;
  size_t st;
  const size_t base = col_addr<AX>(d, col, st);
  C u[R1];
#pragma unroll
  for (int j = 0; j < R1; ++j) u[j] = s[base + (size_t)(t0 + j * T) * st];
  Rad<L>::template run_regs_to<INV>(u, lds + xl0, TX, t0, tw, [&](int z, C v) {
    if (live) s[base + (size_t)z * st] = v;
  });
}

// z convolution: [forward z] -> (x kern_{b,o} -> inverse z -> out_{b,o}) for o < NOUT.
// One LDS tile of TX columns.  Thread (column xl0, t0) owns the elements
// t0 + j T of its column, which are the operands of the first radix pass in
// both directions: the forward transform's first pass reads them straight
// from global memory, the forward spectrum is kept in registers (one value per
// operand) and re-multiplied for each kernel straight into the first inverse
// pass, and the last inverse pass stores to global memory; only the passes in
// between go through LDS.
template <int L, int NOUT, typename R>
__global__ void __launch_bounds__(ColTile<L>::NT)
    k_zconv(ZConvArgs<R> a, Dims d, long long ncol, const cplx_t<R>* __restrict__ tw) {
  using C = cplx_t<R>;
  constexpr int TX = ColTile<L>::TX, NT = ColTile<L>::NT, T = Rad<L>::T, R1 = L / T;
  static_assert(NT == T * TX, "one thread per (column, first-pass butterfly)");
  __shared__ C src[L * TX];
  const int b = blockIdx.y;
  const C* in = a.in[b];
  const size_t stride = (size_t)d.nxh * d.ny;
  const long long c0 = (long long)blockIdx.x * TX;
  const int xl0 = threadIdx.x % TX, t0 = threadIdx.x / TX;
  const long long col = min(c0 + xl0, ncol - 1);  // clamped columns are never stored
  const bool live = c0 + xl0 < ncol;
  C keep[R1];
#pragma unroll
  for (int j = 0; j < R1; ++j) keep[j] = in[col + (size_t)(t0 + j * T) * stride];
  if (a.do_fwd) {
    Rad<L>::template run_from_regs<false>(keep, src + xl0, TX, t0, tw);
    if (a.store[b]) {
      for (int e = threadIdx.x; e < TX * L; e += NT) {
        const int xl = e % TX, j = e / TX;
        if (c0 + xl < ncol) a.store[b][c0 + xl + j * stride] = src[e];
      }
    }
#pragma unroll
    for (int j = 0; j < R1; ++j) keep[j] = src[xl0 + (t0 + j * T) * TX];
  }
#pragma unroll 1
  for (int o = 0; o < NOUT; ++o) {
    const R* kern = a.kern[b * NOUT + o];
    C u[R1];
#pragma unroll
    for (int j = 0; j < R1; ++j) u[j] = cscale(keep[j], kern[col + (size_t)(t0 + j * T) * stride]);
    if (o > 0 || a.do_fwd) __syncthreads();  // earlier readers of the tile are done
    C* out = a.out[b * NOUT + o];
    Rad<L>::template run_regs_to<true>(u, src + xl0, TX, t0, tw, [&](int z, C v) {
      if (live) out[col + (size_t)z * stride] = v;
    });
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int M>
static constexpr int xrows() {
  return 256 / Rad<M>::T > 64 ? 64 : 256 / Rad<M>::T;
}

static int dispatch_len(int L) {
  switch (L) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256: return 1;
    default: return 0;
  }
}

#define HEDFT_LEN_SWITCH(LEN, ...)           \
  switch (LEN) {                              \
    case 16: { constexpr int L_ = 16; __VA_ARGS__; } break;   \
    case 24: { constexpr int L_ = 24; __VA_ARGS__; } break;   \
    case 32: { constexpr int L_ = 32; __VA_ARGS__; } break;   \
    case 48: { constexpr int L_ = 48; __VA_ARGS__; } break;   \
    case 64: { constexpr int L_ = 64; __VA_ARGS__; } break;   \
    case 96: { constexpr int L_ = 96; __VA_ARGS__; } break;   \
    case 128: { constexpr int L_ = 128; __VA_ARGS__; } break; \
    case 192: { constexpr int L_ = 192; __VA_ARGS__; } break; \
    case 256: { constexpr int L_ = 256; __VA_ARGS__; } break; \
    default: return HEDFT_EGRID;              \
  }

// square planes small enough for one workgroup's LDS take the fused x+y path
#define HEDFT_PLANE_SWITCH(NXV, NYV, ...)                                       \
  if (NXV == NYV) {                                                              \
    switch (NXV) {                                                               \
      case 32: { constexpr int PX_ = 32, PY_ = 32; __VA_ARGS__; } return HEDFT_OK;   \
      case 48: { constexpr int PX_ = 48, PY_ = 48; __VA_ARGS__; } return HEDFT_OK;   \
      case 64: { constexpr int PX_ = 64, PY_ = 64; __VA_ARGS__; } return HEDFT_OK;   \
      case 96: { constexpr int PX_ = 96, PY_ = 96; __VA_ARGS__; } return HEDFT_OK;   \
      case 128: { constexpr int PX_ = 128, PY_ = 128; __VA_ARGS__; } return HEDFT_OK; \
      default: break;                                                            \
    }                                                                            \
  }

template <typename R>
static inline const cplx_t<R>* twp(const hedft_ctx* c, size_t off) {
  if constexpr (sizeof(R) == 8) return reinterpret_cast<const double2*>(c->tw) + off;
  else return reinterpret_cast<const float2*>(c->twf) + off;
}

static bool g_plane_attr_set[8];

template <int N, typename R>
static hipError_t plane_attrs() {
  hipError_t e = hipFuncSetAttribute((const void*)k_xyfwd<N, N, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)Plane<N, N, R>::LDS);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_xyinv<N, N, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)Plane<N, N, R>::LDS);
  if (e != hipSuccess || N != 128) return e;
  e = hipFuncSetAttribute((const void*)k_xyfwd128<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)Plane<N, N, R>::LDS);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_xyinv128<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)Plane<N, N, R>::LDS);
}

// the 128 x 128 plane kernels with fewer LDS round trips, for double storage
// (HEDFT_PLANE128=0 in the environment selects the generic ones).  With float
// storage the generic kernels stay: their 67 KB planes and 61-64 VGPRs let two
// workgroups share a CU, which the register-resident passes (84-86 VGPRs) do
// not, and both measure the same at 128^3.
template <typename R>
static bool plane128() {
  static const int enabled = [] {
    const char* e = std::getenv("HEDFT_PLANE128");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  return enabled && sizeof(R) == 8;
}

int fft_setup(hedft_ctx* c) {
  const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz, M = nx / 2;
  if (!dispatch_len(M) || !dispatch_len(ny) || !dispatch_len(nz) || (nx & 1)) return HEDFT_EGRID;
  std::vector<double2> h;
  auto add = [&](int n, int count) {
    size_t off = h.size();
    for (int m = 0; m < count; ++m) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * m / n;
      h.push_back(make_double2((double)cosl(a), (double)sinl(a)));
    }
    return off;
  };
  c->tw_x = add(M, M);
  c->tw_xpost = add(nx, M + 1);
  c->tw_y = add(ny, ny);
  c->tw_z = add(nz, nz);
  std::vector<float2> hf(h.size());
  for (size_t i = 0; i < h.size(); ++i) hf[i] = make_float2((float)h[i].x, (float)h[i].y);
  HEDFT_CHECK(hipMalloc(&c->tw, h.size() * sizeof(double2)));
  HEDFT_CHECK(hipMemcpy(c->tw, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
  HEDFT_CHECK(hipMalloc(&c->twf, hf.size() * sizeof(float2)));
  HEDFT_CHECK(hipMemcpy(c->twf, hf.data(), hf.size() * sizeof(float2), hipMemcpyHostToDevice));
  // plane kernels use more than 64 KB of dynamic LDS
  const int sizes[5] = {32, 48, 64, 96, 128};
  for (int k = 0; k < 5; ++k) {
    if (g_plane_attr_set[k]) continue;
    hipError_t e = hipSuccess;
    switch (sizes[k]) {
#define HEDFT_SETATTR(N)                                 \
  case N:                                                \
    e = plane_attrs<N, double>();                        \
    if (e == hipSuccess) e = plane_attrs<N, float>();    \
    break;
      HEDFT_SETATTR(32) HEDFT_SETATTR(48) HEDFT_SETATTR(64) HEDFT_SETATTR(96) HEDFT_SETATTR(128)
#undef HEDFT_SETATTR
    }
    if (e != hipSuccess) return HEDFT_EHIP;
    g_plane_attr_set[k] = true;
  }
  return HEDFT_OK;
}

template <typename R>
static int launch_xr2c(hedft_ctx* c, const XInBatch<R>& a, int nb) {
  const int M = c->d.nx / 2;
  const int nrows = c->d.ny * c->d.nz;
  HEDFT_LEN_SWITCH(M, {
    constexpr int ROWS = xrows<L_>();
    dim3 grid((nrows + ROWS - 1) / ROWS, nb);
    hipLaunchKernelGGL((k_xr2c<L_, ROWS, R>), grid, dim3(Rad<L_>::T * ROWS), 0, c->stream, a, c->d,
                       twp<R>(c, c->tw_x), twp<R>(c, c->tw_xpost));
  });
  return HEDFT_OK;
}

template <typename R>
static int launch_xc2r(hedft_ctx* c, const XOutBatch<R>& a, int nb) {
  const int M = c->d.nx / 2;
  const int nrows = c->d.ny * c->d.nz;
  HEDFT_LEN_SWITCH(M, {
    constexpr int ROWS = xrows<L_>();
    dim3 grid((nrows + ROWS - 1) / ROWS, nb);
    hipLaunchKernelGGL((k_xc2r<L_, ROWS, R>), grid, dim3(Rad<L_>::T * ROWS), 0, c->stream, a, c->d,
                       twp<R>(c, c->tw_x), twp<R>(c, c->tw_xpost), c->renor);
  });
  return HEDFT_OK;
}

template <int AX, bool INV, typename R>
static int launch_col(hedft_ctx* c, const SpecBatch<R>& a, int nb) {
  const int L = (AX == 1) ? c->d.ny : c->d.nz;
  const size_t off = (AX == 1) ? c->tw_y : c->tw_z;
  const long long ncol = (long long)c->d.nxh * ((AX == 1) ? c->d.nz : c->d.ny);
  HEDFT_LEN_SWITCH(L, {
    constexpr int TX = ColTile<L_>::TX;
    dim3 grid((unsigned)((ncol + TX - 1) / TX), nb);
    hipLaunchKernelGGL((k_colfft<L_, AX, INV, R>), grid, dim3(ColTile<L_>::NT), 0, c->stream, a, c->d, ncol,
                       twp<R>(c, off));
  });
  return HEDFT_OK;
}

template <typename R>
int fwd_xy(hedft_ctx* c, int nb, const XInT<R>* in, cplx_t<R>* const* spec) {
  if (nb < 1 || nb > MAXB) return HEDFT_EINVAL;
  XInBatch<R> a{};
  SpecBatch<R> s{};
  for (int b = 0; b < nb; ++b) {
    a.in[b] = in[b];
    a.out[b] = spec[b];
    s.p[b] = spec[b];
  }
  if (nb > c->xy_split_nb[0] && c->d.nx == 128 && c->d.ny == 128 && plane128<R>()) {
    constexpr size_t lds128 = Plane<128, 128, R>::LDS;
    hipLaunchKernelGGL((k_xyfwd128<R>), dim3(c->d.nz, nb), dim3(1024), lds128, c->stream, a,
                       twp<R>(c, c->tw_x), twp<R>(c, c->tw_xpost), twp<R>(c, c->tw_y));
    return HEDFT_OK;
  }
  if (nb > c->xy_split_nb[0]) HEDFT_PLANE_SWITCH(c->d.nx, c->d.ny, {
    using PL_ = Plane<PX_, PY_, R>;
    constexpr size_t lds_ = PL_::LDS;
    constexpr int nt_ = PL_::NT;
    hipLaunchKernelGGL((k_xyfwd<PX_, PY_, R>), dim3(c->d.nz, nb), dim3(nt_), lds_, c->stream, a, c->d,
                       twp<R>(c, c->tw_x), twp<R>(c, c->tw_xpost), twp<R>(c, c->tw_y));
  });
  HEDFT_TRY(launch_xr2c<R>(c, a, nb));
  return launch_col<1, false, R>(c, s, nb);
}

template <typename R>
int zconv(hedft_ctx* c, int nin, cplx_t<R>* const* in, int nout, const R* const* kern,
          cplx_t<R>* const* out, int do_fwd, cplx_t<R>* const* store_fwd) {
  if (nin < 1 || nin > MAXB || nout < 1 || nout > 3) return HEDFT_EINVAL;
  ZConvArgs<R> a{};
  for (int i = 0; i < nin; ++i) {
    a.in[i] = in[i];
    a.store[i] = store_fwd ? store_fwd[i] : nullptr;
    for (int o = 0; o < nout; ++o) {
      a.out[i * nout + o] = out[i * nout + o];
      a.kern[i * nout + o] = kern[i * nout + o];
    }
  }
  a.do_fwd = do_fwd;
  const int L = c->d.nz;
  const long long ncol = (long long)c->d.nxh * c->d.ny;
  HEDFT_LEN_SWITCH(L, {
    constexpr int TX = ColTile<L_>::TX;
    dim3 grid((unsigned)((ncol + TX - 1) / TX), nin);
    dim3 blk(ColTile<L_>::NT);
    const cplx_t<R>* tw = twp<R>(c, c->tw_z);
    if (nout == 1) hipLaunchKernelGGL((k_zconv<L_, 1, R>), grid, blk, 0, c->stream, a, c->d, ncol, tw);
    else if (nout == 2) hipLaunchKernelGGL((k_zconv<L_, 2, R>), grid, blk, 0, c->stream, a, c->d, ncol, tw);
    else hipLaunchKernelGGL((k_zconv<L_, 3, R>), grid, blk, 0, c->stream, a, c->d, ncol, tw);
  });
  return HEDFT_OK;
}

template <typename R>
int inv_yx(hedft_ctx* c, int nb, cplx_t<R>* const* spec, const XOutT<R>* out) {
  if (nb < 1 || nb > MAXB) return HEDFT_EINVAL;
  SpecBatch<R> s{};
  XOutBatch<R> a{};
  for (int b = 0; b < nb; ++b) {
    s.p[b] = spec[b];
    a.in[b] = spec[b];
    a.out[b] = out[b];
  }
  if (nb > c->xy_split_nb[1] && c->d.nx == 128 && c->d.ny == 128 && plane128<R>()) {
    constexpr size_t lds128 = Plane<128, 128, R>::LDS;
    hipLaunchKernelGGL((k_xyinv128<R>), dim3(c->d.nz, nb), dim3(1024), lds128, c->stream, a,
                       twp<R>(c, c->tw_x), twp<R>(c, c->tw_xpost), twp<R>(c, c->tw_y), c->renor);
    return HEDFT_OK;
  }
  if (nb > c->xy_split_nb[1]) HEDFT_PLANE_SWITCH(c->d.nx, c->d.ny, {
    using PL_ = Plane<PX_, PY_, R>;
    constexpr size_t lds_ = PL_::LDS;
    constexpr int nt_ = PL_::NT;
    hipLaunchKernelGGL((k_xyinv<PX_, PY_, R>), dim3(c->d.nz, nb), dim3(nt_), lds_, c->stream, a, c->d,
                       twp<R>(c, c->tw_x), twp<R>(c, c->tw_xpost), twp<R>(c, c->tw_y), c->renor);
  });
  HEDFT_TRY((launch_col<1, true, R>(c, s, nb)));
  return launch_xc2r<R>(c, a, nb);
}

template int fwd_xy<double>(hedft_ctx*, int, const XInT<double>*, double2* const*);
template int fwd_xy<float>(hedft_ctx*, int, const XInT<float>*, float2* const*);
template int zconv<double>(hedft_ctx*, int, double2* const*, int, const double* const*, double2* const*, int,
                           double2* const*);
template int zconv<float>(hedft_ctx*, int, float2* const*, int, const float* const*, float2* const*, int,
                          float2* const*);
template int inv_yx<double>(hedft_ctx*, int, double2* const*, const XOutT<double>*);
template int inv_yx<float>(hedft_ctx*, int, float2* const*, const XOutT<float>*);

int fftfw(hedft_ctx* c, const double* in, double2* out) {
  XIn xi{};
  xi.op = XIN_COPY;
  xi.a[0] = in;
  HEDFT_TRY(fwd_xy<double>(c, 1, &xi, &out));
  SpecBatch<double> s{};
  s.p[0] = out;
  return launch_col<2, false, double>(c, s, 1);
}

int fftbk(hedft_ctx* c, double2* in, double* out) {
  SpecBatch<double> s{};
  s.p[0] = in;
  HEDFT_TRY((launch_col<2, true, double>(c, s, 1)));
  XOut xo{};
  xo.op = XOUT_STORE;
  xo.out = out;
  return inv_yx<double>(c, 1, &in, &xo);
}

}  // namespace hedft
