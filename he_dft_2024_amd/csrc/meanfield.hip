// meanfield.hip — the Orsay-Trento mean field: spectral kernel tables
// (fforma, initcg, kalfs), the convolution stages of poten and term_alfa, the
// quantum-impurity potentials, and the statics refresh of the driver.  Two
// streams: independent branches run on the side stream (fork / join).
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
void kernel_tables(const hedft_ctx* c, std::vector<double>& fv, std::vector<double>& wc,
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
// convolution and one inverse of the impurity density
// F^-1[F(|psix_k|^2) kern] into out (stored, or added with acc)
template <typename R>
static int qconv(hedft_ctx* c, double* out, int k, const double* kern, bool acc) {
  XInT<R> xa{};
  xa.op = XIN_ABS2;
  xa.z = (const double2*)c->qi[k].psix;
  cplx_t<R>* sx = spc<R>(c->specq[0]);
  HEDFT_TRY(fwd_xy<R>(c, 1, &xa, &sx));
  const R* ku = fld<R>(const_cast<double*>(kern));
  HEDFT_TRY(zconv<R>(c, 1, &sx, 1, &ku, &sx, 1, nullptr));
  XOutT<R> xo = acc ? xo_acc<R>(out) : xo_store<R>(out);
  return inv_yx<R>(c, 1, &sx, &xo);
}
template <typename R>
static int potx4_conv(hedft_ctx* c, double* out, int k = 0, bool acc = false) {
  return qconv<R>(c, out, k, c->qi[k].uimpk, acc);
}

// Several quantum impurities (potenimp.f90:13-39), from the x+y spectrum of
// den in sden: for each impurity i, potx4 += F^-1[F(|psix_i|^2) uimp_k_i],
// upotx_i = F^-1[fden uimp_k_i] + sum_{j != i} F^-1[F(|psix_j|^2)
// uimp_k_k(i,j)], each term convolved and transformed back on its own and
// summed in real space in the reference's order (specq[0] holds |psix_j|^2's
// x+y spectrum, specq[1] a term's spectrum)
template <typename R>
static int potenimp_multi(hedft_ctx* c, cplx_t<R>* sden) {
  cplx_t<R>* sx = spc<R>(c->specq[0]);
  cplx_t<R>* st = spc<R>(c->specq[1]);
  auto fwd_denx = [&](int j) {
    XInT<R> xa{};
    xa.op = XIN_ABS2;
    xa.z = (const double2*)c->qi[j].psix;
    return fwd_xy<R>(c, 1, &xa, &sx);
  };
  auto term = [&](cplx_t<R>* sxy, const double* kern, double* out, bool acc) {
    const R* ku = fld<R>(const_cast<double*>(kern));
    HEDFT_TRY(zconv<R>(c, 1, &sxy, 1, &ku, &st, 1, nullptr));
    XOutT<R> xo = acc ? xo_acc<R>(out) : xo_store<R>(out);
    return inv_yx<R>(c, 1, &st, &xo);
  };
  for (int i = 0; i < c->quantum; ++i) {
    HEDFT_TRY(fwd_denx(i));
    HEDFT_TRY(term(sx, c->qi[i].uimpk, c->potx4, i > 0));
    HEDFT_TRY(term(sden, c->qi[i].uimpk, c->qi[i].upotx, false));
    for (int j = 0; j < c->quantum; ++j) {
      if (j == i || !c->uimpkk[i][j]) continue;
      HEDFT_TRY(fwd_denx(j));
      HEDFT_TRY(term(sx, c->uimpkk[i][j], c->qi[i].upotx, true));
    }
  }
  return HEDFT_OK;
}

// potenimp of Dynamics_quantum_impurities (potenimp.f90:13-39) on its own:
// upotx = F^-1[F(den) uimp_k], potx4 (one impurity), or every impurity's
// terms
template <typename R>
static int potenimp_q(hedft_ctx* c) {
  XInT<R> xi{};
  xi.op = XIN_COPY;
  xi.a[0] = fld<R>(c->den);
  cplx_t<R>* s0 = spc<R>(c->spec[0]);
  HEDFT_TRY(fwd_xy<R>(c, 1, &xi, &s0));
  if (c->quantum > 1) return potenimp_multi<R>(c, s0);
  const R* ku = fld<R>(c->qi[0].uimpk);
  cplx_t<R>* s1 = spc<R>(c->specq[1]);
  HEDFT_TRY(zconv<R>(c, 1, &s0, 1, &ku, &s1, 1, nullptr));
  XOutT<R> xo = xo_store<R>(c->qi[0].upotx);
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
// convolved on the side branch and added in the pot4 assembly.  With several
// quantum impurities their terms (potenimp_multi) follow the den convolutions
// on the main stream, while den's x+y spectrum is still in spec[0].
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
    if (c->quantum == 1) HEDFT_TRY(potx4_conv<R>(c, c->potx4));
    if (mid) {
      HEDFT_CHECK(hipEventRecord(c->ev_mid, c->side_stream));
      if (derden && want_force) HEDFT_TRY(derden_force_finalize(c, want_force));
      flush_side_errors(c);
    }
    if (with_lap && !late && !on_main) {
      HEDFT_TRY(lap_xy(c, c->psi, 1));
      for (int k = 0; k < c->quantum; ++k) HEDFT_TRY(lap_xy(c, c->qi[k].psix, 1, c->qi[k].lapx));
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
  if (c->quantum == 1) {
    kk[nk] = fld<R>(c->qi[0].uimpk);
    so[nk] = spc<R>(c->specq[1]);
    xo[nk++] = xo_store<R>(c->qi[0].upotx);
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
      if (xo[k].out == fld<R>(c->delj4) || (c->quantum == 1 && xo[k].out == fld<R>(c->qi[0].upotx))) {
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
  if (c->quantum > 1) HEDFT_TRY(potenimp_multi<R>(c, s0));
  if (on_main) {
    HEDFT_TRY(lap_xy(c, c->psi, 1));
    for (int k = 0; k < c->quantum; ++k) HEDFT_TRY(lap_xy(c, c->qi[k].psix, 1, c->qi[k].lapx));
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
      for (int k = 0; k < c->quantum; ++k) HEDFT_TRY(lap_xy(c, c->qi[k].psix, 1, c->qi[k].lapx));
    }
  }
  if (with_lap) c->lap_fresh = 1;
  HEDFT_TRY(stage_b<R>(c));
  if (c->p.core_otc) HEDFT_TRY(stage_c<R>(c));
  else if (late || mid) HEDFT_TRY(join(c));
  return HEDFT_OK;
}

int poten_tddft(hedft_ctx* c, int want_force, int with_lap) {
  return c->mixed ? poten_tddft_t<float>(c, want_force, with_lap) : poten_tddft_t<double>(c, want_force, with_lap);
}

// He_statics/poten.f90:5-221: fden, dencg, denalf and grad(den) as left by the caller
// (with an impurity wave function also fdenx: upotx = F^-1[fden vq], potx4 =
// F^-1[fdenx vq] enter the pot4 assembly, poten.f90:94-107)
int poten_statics(hedft_ctx* c) {
  const double* kf = c->fvlj4;
  double2* s2 = c->spec[2];
  double2* fd = c->fden;
  HEDFT_TRY(zconv(c, 1, &fd, 1, &kf, &s2, 0, nullptr));
  XOut xo = xo_store<double>(c->delj4);
  HEDFT_TRY(inv_yx(c, 1, &s2, &xo));
  if (c->quantum) {
    double2* in[2] = {c->fden, c->specq[0]};
    const double* kq[2] = {c->qi[0].uimpk, c->qi[0].uimpk};
    double2* so[2] = {c->spec[0], c->spec[1]};
    HEDFT_TRY(zconv(c, 2, in, 1, kq, so, 0, nullptr));
    XOut xq[2] = {xo_store<double>(c->qi[0].upotx), xo_store<double>(c->potx4)};
    HEDFT_TRY(inv_yx(c, 2, so, xq));
  }
  HEDFT_TRY(stage_b<double>(c));
  if (c->p.core_otc) HEDFT_TRY(stage_c<double>(c));
  return HEDFT_OK;
}

// fdenx = F(psix^2) of the statics impurity (BCN4HeDFT.f90:788,1130)
static int statics_fdenx(hedft_ctx* c) {
  hipLaunchKernelGGL(k_sq<double>, dim3(EW_BLOCKS), dim3(256), 0, c->stream, c->d.n, (const double*)c->qi[0].psix, c->work2);
  return fftfw(c, c->work2, c->specq[0]);
}

// driver refresh after a statics step (BCN4HeDFT.f90:1131-1149): derden, fden, dencg, denalf
// (and fdenx with an impurity wave function, BCN4HeDFT.f90:1128-1130)
int statics_refresh(hedft_ctx* c) {
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

int alpha_term_run(hedft_ctx* c) { return c->mixed ? alpha_term<float>(c) : alpha_term<double>(c); }
int potenimp_quantum(hedft_ctx* c) { return c->mixed ? potenimp_q<float>(c) : potenimp_q<double>(c); }
int qconv_run(hedft_ctx* c, double* out, int k, const double* kern, int acc) {
  return c->mixed ? qconv<float>(c, out, k, kern, acc != 0) : qconv<double>(c, out, k, kern, acc != 0);
}

}  // namespace hedft
