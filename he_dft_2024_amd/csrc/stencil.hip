// stencil.hip — derivnD coefficient set-up and the non-step stencil passes.
//
// Coefficients: the Bickley 13-point matrices that newder.f returns with
// ICON=1 (rows = differentiation node, table mates.f:172), generated here with
// Fornberg's recurrence (Math. Comp. 51 (1988) 699) in long double; the
// boundary rows are then rewritten for the boundary mode icon as redef.f:1-324
// does (13: zero outside the box; 8/18: (anti)periodic wrap; 24 = 20 + 4:
// even reflection about both end points; 0: one-sided rows) and divided by
// h^id (init_deriv_parallel.f90:399-414).
#include <cmath>
#include <cstring>

#include "stencil_kernel.h"

namespace hedft {

static void bickley_matrix(int id, double cd[13][13]) {
  for (int i = 0; i < 13; ++i) {
    long double w[13][3];
    std::memset(w, 0, sizeof w);
    const long double z = (long double)i;
    long double c1 = 1.0L, c4 = 0.0L - z;
    w[0][0] = 1.0L;
    for (int a = 1; a < 13; ++a) {
      const int mn = a < id ? a : id;
      long double c2 = 1.0L, c5 = c4;
      c4 = (long double)a - z;
      for (int b = 0; b < a; ++b) {
        const long double c3 = (long double)(a - b);
        c2 *= c3;
        if (b == a - 1) {
          for (int k = mn; k >= 1; --k) w[a][k] = c1 * (k * w[a - 1][k - 1] - c5 * w[a - 1][k]) / c2;
          w[a][0] = -c1 * c5 * w[a - 1][0] / c2;
        }
        for (int k = mn; k >= 1; --k) w[b][k] = (c4 * w[b][k] - k * w[b][k - 1]) / c3;
        w[b][0] = c4 * w[b][0] / c3;
      }
      c1 = c2;
    }
    for (int j = 0; j < 13; ++j) cd[i][j] = (double)w[j][id];
  }
}

// Boundary-row rewrite for one mode step (the cases of redef.f on this path).
static void redef_rows(double cd[13][13], int icon) {
  const int K = 13, ko2 = 6;
  while (icon != 0) {
    int next = 0;
    if (icon == 4 || icon == 5) {  // even / odd reflection about the first point
      const int ip = (icon == 4) ? 1 : -1;
      double c0[13];
      for (int j = 0; j < K; ++j) c0[j] = cd[ko2][j];
      for (int i = 1; i <= ko2; ++i) {
        int j0 = ko2 + 2 - i, j1 = j0;
        cd[i - 1][0] = c0[j0 - 1];
        for (int j = 2; j <= K; ++j) {
          ++j0;
          --j1;
          cd[i - 1][j - 1] = (j0 <= K) ? ((j1 >= 1) ? c0[j0 - 1] + ip * c0[j1 - 1] : c0[j0 - 1]) : 0.0;
        }
      }
    } else if (icon == 12) {  // zero beyond the last point
      for (int i = ko2 + 2; i <= K; ++i) {
        for (int j = K; j >= 2; --j) cd[i - 1][j - 1] = cd[i - 2][j - 2];
        cd[i - 1][0] = 0.0;
      }
    } else if (icon == 13) {  // zero before the first point, then 12
      for (int i = ko2; i >= 1; --i) {
        for (int j = 1; j <= K - 1; ++j) cd[i - 1][j - 1] = cd[i][j];
        cd[i - 1][K - 1] = 0.0;
      }
      next = 12;
    } else if (icon == 20 || icon == 24 || icon == 25) {  // even reflection about the last point
      double c0[13];
      for (int j = 0; j < K; ++j) c0[j] = cd[ko2][j];
      for (int i = K; i >= ko2 + 2; --i) {
        int j0 = ko2 + 1 - i + K, j1 = j0;
        cd[i - 1][K - 1] = c0[j0 - 1];
        for (int j = K - 1; j >= 1; --j) {
          --j0;
          ++j1;
          cd[i - 1][j - 1] = (j0 >= 1) ? ((j1 <= K) ? c0[j0 - 1] + c0[j1 - 1] : c0[j0 - 1]) : 0.0;
        }
      }
      next = (icon == 24) ? 4 : (icon == 25) ? 5 : 0;
    }
    icon = next;
  }
}

static StencilCoef make_coef(int id, int icon, double h) {
  StencilCoef cf;
  double cd[13][13];
  bickley_matrix(id, cd);
  if (icon != 0 && icon != 8 && icon != 18) redef_rows(cd, icon);
  const double sto = 1.0 / std::pow(h, id);
  for (int j = 0; j < 13; ++j) cf.c[j] = cd[6][j] * sto;
  for (int r = 0; r < 6; ++r)
    for (int j = 0; j < 13; ++j) {
      cf.rows[r][j] = cd[r][j] * sto;
      cf.rows[6 + r][j] = cd[7 + r][j] * sto;
    }
  cf.mode = (icon == 8) ? 1 : (icon == 18) ? 2 : 0;
  return cf;
}

void stencil_setup(hedft_ctx* c) {
  const double h[3] = {c->hx, c->hy, c->hz};
  for (int id = 1; id <= 2; ++id)
    for (int a = 0; a < 3; ++a) {
      c->coef[id - 1][a] = make_coef(id, c->p.icon, h[a]);
      c->coef0[id - 1][a] = make_coef(id, 0, h[a]);
    }
}

template <typename V>
static int deriv_plain_t(hedft_ctx* c, int id, int axis, const V* in, V* out) {
  const StencilCoef& cf = c->coef[id - 1][axis];
  InPlain<V> ip{in};
  OutStore<V> op{out, nullptr};
  switch (axis) {
    case 0: launch_pencil<V, 0>(c->stream, c->d, cf, ip, op); break;
    case 1: launch_pencil<V, 1>(c->stream, c->d, cf, ip, op); break;
    default: launch_pencil<V, 2>(c->stream, c->d, cf, ip, op); break;
  }
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

int deriv_plain(hedft_ctx* c, int id, int axis, const void* in, void* out, int is_complex) {
  if (c->d.nx < 13 || c->d.ny < 13 || c->d.nz < 13) return HEDFT_EGRID;
  if (is_complex)
    return deriv_plain_t<double2>(c, id, axis, (const double2*)in, (double2*)out);
  return deriv_plain_t<double>(c, id, axis, (const double*)in, (double*)out);
}

// grad(den): x and y in one fused pass, z in a second pass that also
// accumulates the He part of the impurity force F_k = -sum(grad(den) U_k) dxyz
// (potenimp.f90:49-52), evaluating U_k once per point.  R is the storage type
// of the density fields (double, or float on the mixed path).
// NI impurities with force accumulators: 0 (the gradient alone), 1 (the usual
// single impurity, fully unrolled) or HEDFT_MAX_IMP (any count, run-time loop)
template <typename R, int NI>
struct OutGradZForce {
  static constexpr int NR = 3 * NI;
  R* out;
  const R *gx, *gy;
  double* partials;
  ImpDev im;
  __device__ __forceinline__ void store(size_t i, int ix, int iy, int iz, double v, double* red) const {
    out[i] = (R)v;
    if constexpr (NI > 0) {
      double xx, yy, zz;
      imp_xyz(im, ix, iy, iz, xx, yy, zz);
      const double a = gx[i], b = gy[i];
      const int nk = (NI == 1) ? 1 : im.nimp;
#pragma unroll
      for (int k = 0; k < NI; ++k) {
        if (k >= nk) break;
        const double u = imp_u_k(im, k, xx, yy, zz, false);
        red[3 * k] += a * u;
        red[3 * k + 1] += b * u;
        red[3 * k + 2] += v * u;
      }
    }
  }
};


template <typename R>
static int derden_force_t(hedft_ctx* c, int want_force, int update_aimp, int finalize) {
  want_force = want_force && c->nimp > 0;
  R* den = reinterpret_cast<R*>(c->den);
  R* dx = reinterpret_cast<R*>(c->dxden);
  R* dy = reinterpret_cast<R*>(c->dyden);
  R* dz = reinterpret_cast<R*>(c->dzden);
  InPlain<R> ip{den};
  launch_pencil_xy<R, true>(c->stream, c->d, c->coef[0][0], c->coef[0][1], ip, OutTwo<R>{dx, dy, nullptr});
  if (want_force && c->nimp == 1)
    launch_pencil<R, 2>(c->stream, c->d, c->coef[0][2], ip,
                        OutGradZForce<R, 1>{dz, dx, dy, force_partials(c), make_impdev(c)});
  else if (want_force)
    launch_pencil<R, 2>(c->stream, c->d, c->coef[0][2], ip,
                        OutGradZForce<R, HEDFT_MAX_IMP>{dz, dx, dy, force_partials(c), make_impdev(c)});
  else
    launch_pencil<R, 2>(c->stream, c->d, c->coef[0][2], ip,
                        OutGradZForce<R, 0>{dz, dx, dy, nullptr, make_impdev(c)});
  if (hipGetLastError() != hipSuccess) return HEDFT_EHIP;
  return (want_force && finalize) ? derden_force_finalize(c, update_aimp) : HEDFT_OK;
}

int derden_force(hedft_ctx* c, int want_force, int update_aimp, int finalize) {
  return c->mixed ? derden_force_t<float>(c, want_force, update_aimp, finalize)
                  : derden_force_t<double>(c, want_force, update_aimp, finalize);
}

int derden_force_finalize(hedft_ctx* c, int update_aimp) {
  if (c->nimp <= 0) return HEDFT_OK;
  const int nr = (c->nimp == 1) ? 3 : 3 * HEDFT_MAX_IMP;  // accumulators per block of the z pass
  const size_t nb = c->mixed ? pencil_geom<float>(c->d, 2).nblocks : pencil_geom<double>(c->d, 2).nblocks;
  return force_finalize(c, nb, update_aimp, nr);
}

// z pass of the alpha_s divergence with the mean-field assembly:
// ualphas = (conv/den0s + sum_a d_a(f int_a)) h2o2m4 alphas   (term_alfa.f90:80-98)
// pot4 = delj4 + dencg^2 (cp4/2 + cpp4/3 dencg) + corr [+ uext | potx4] + ualphas
// (He_statics/poten.f90:73-91, Dynamics_TDDFT/poten.f90:194-211,
// Dynamics_quantum_impurities/poten.f90:116-138)
template <typename R>
struct OutDivPot4 {
  static constexpr int NR = 0;
  double* partials;
  const R *acc, *conv, *delj4, *dencg, *corr, *ext;
  R *ualphas, *pot4;
  double den0s, halpha, a0, a1;
  int lext;
  __device__ __forceinline__ void store(size_t i, int, int, int, double v, double*) const {
    const AlphaPot4 r = alpha_pot4((double)conv[i], den0s, (double)acc[i] + v, halpha, (double)delj4[i],
                                   (double)dencg[i], (double)corr[i], a0, a1, lext != 0,
                                   lext ? (double)ext[i] : 0.0);
    ualphas[i] = (R)r.ua;
    pot4[i] = (R)r.pot4;
  }
};

// x + y part of the divergence sum_a d_a(f int_a), f = 1 - denalf/den0s, into work1
template <typename R>
static int alpha_div_xy_t(hedft_ctx* c) {
  auto F = [](double* p) { return reinterpret_cast<R*>(p); };
  InFInt2<R> ixy{F(c->denalf), F(c->intx), F(c->inty), c->p.den0s};
  launch_pencil_xy<R, false>(c->stream, c->d, c->coef[0][0], c->coef[0][1], ixy,
                             OutSumXY<R>{F(c->work1), nullptr});
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

int alpha_div_xy(hedft_ctx* c) { return c->mixed ? alpha_div_xy_t<float>(c) : alpha_div_xy_t<double>(c); }

// z part of the divergence, fused with the ualphas and pot4 assembly
template <typename R>
static int alpha_div_z_pot4_t(hedft_ctx* c) {
  auto F = [](double* p) { return reinterpret_cast<R*>(p); };
  const double den0s = c->p.den0s;
  InFInt<R> iz{F(c->denalf), F(c->intz), den0s};
  OutDivPot4<R> op;
  op.partials = nullptr;
  op.acc = F(c->work1);
  op.conv = F(c->ualph_conv);
  op.delj4 = F(c->delj4);
  op.dencg = F(c->dencg);
  op.corr = F(c->sto1);
  op.ext = F(pot4_ext(c));
  op.ualphas = F(c->ualphas);
  op.pot4 = F(c->pot4);
  op.den0s = den0s;
  op.halpha = c->p.h2o2m4 * c->p.alphas;
  op.a0 = c->p.cp4 / 2.0;
  op.a1 = c->p.cpp4 / 3.0;
  op.lext = pot4_ext(c) != nullptr;
  launch_pencil<R, 2>(c->stream, c->d, c->coef[0][2], iz, op);
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

int alpha_div_z_pot4(hedft_ctx* c) {
  return c->mixed ? alpha_div_z_pot4_t<float>(c) : alpha_div_z_pot4_t<double>(c);
}

int div25_ok(const hedft_ctx* c) { return lap25_geom(c->d, 16).nblocks > 0; }

// the divergence in one 2.5-D pass (z chunks of 16 planes), with the same
// assembly as alpha_div_z_pot4
template <typename R>
static int alpha_div25_pot4_t(hedft_ctx* c) {
  auto F = [](double* p) { return reinterpret_cast<R*>(p); };
  Div25Args<R> a;
  a.denalf = F(c->denalf);
  a.ix = F(c->intx);
  a.iy = F(c->inty);
  a.iz = F(c->intz);
  a.conv = F(c->ualph_conv);
  a.delj4 = F(c->delj4);
  a.dencg = F(c->dencg);
  a.corr = F(c->sto1);
  a.ext = F(pot4_ext(c));
  a.ualphas = F(c->ualphas);
  a.pot4 = F(c->pot4);
  a.den0s = c->p.den0s;
  a.halpha = c->p.h2o2m4 * c->p.alphas;
  a.a0 = c->p.cp4 / 2.0;
  a.a1 = c->p.cpp4 / 3.0;
  a.lext = pot4_ext(c) != nullptr;
  launch_div25<R>(c->stream, c->d, c->coef1_dev, a, 16);
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

int alpha_div25_pot4(hedft_ctx* c) {
  if (!div25_ok(c)) return HEDFT_EGRID;
  return c->mixed ? alpha_div25_pot4_t<float>(c) : alpha_div25_pot4_t<double>(c);
}

int lap_xy(hedft_ctx* c, const void* psi, int is_complex, void* out) {
  if (!out) out = c->lap;
  if (is_complex)
    launch_pencil_xy<double2, true>(c->stream, c->d, c->coef[1][0], c->coef[1][1],
                                    InPlain<double2>{(const double2*)psi},
                                    OutSumXY<double2>{(double2*)out, nullptr});
  else
    launch_pencil_xy<double, true>(c->stream, c->d, c->coef[1][0], c->coef[1][1],
                                   InPlain<double>{(const double*)psi}, OutSumXY<double>{(double*)out, nullptr});
  return hipGetLastError() == hipSuccess ? HEDFT_OK : HEDFT_EHIP;
}

}  // namespace hedft
