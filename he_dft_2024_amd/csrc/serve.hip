// serve.hip — continuous-batching request loop over independent TDDFT droplets.
//
// No counterpart in the reference, which runs one droplet per process.  Jobs
// (a droplet + classical or quantum impurity set-up and a number of PC iterations) arrive
// over time; a job is admitted when its device footprint fits the HBM budget,
// its psi and PC history stay resident in its own context until it finishes,
// and the active jobs advance in quanta of `chunk` PC iterations, each on its
// own streams, so that their kernels overlap on the GPU.  A job's arithmetic
// does not depend on which other jobs share the device: every result equals
// the result of running that job alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"

namespace {

constexpr double kPstok = 7.63822291;    // ps <-> K^-1 (DFT4He3d.f90:242)
constexpr double kMpU = 0.02061484;      // proton mass in K^-1 A^-2 (modules.f90: mp_u)

// SplitMix64 uniform in [0,1): the density noise of a job depends on its seed only
struct Rng {
  uint64_t s;
  double uniform() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

// Woods-Saxon droplet rho0/(1+exp((r-R)/0.7)), R = 2.22 N^(1/3), times
// (1 + 1e-3 U), normalised to N (BCN4HeDFT.f90:730-734)
std::vector<double> droplet_density(const hedft_job& j) {
  const int n = j.n;
  const double xmax = 0.5 * n * j.h, R = 2.22 * std::cbrt(j.n4);
  std::vector<double> rho((size_t)n * n * n);
  Rng rng{j.seed};
  double sum = 0.0;
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        const double x = -xmax + j.h * ix, y = -xmax + j.h * iy, z = -xmax + j.h * iz;
        const double r = std::sqrt(x * x + y * y + z * z);
        const double v = 0.0218 / (1.0 + std::exp((r - R) / 0.7)) * (1.0 + 1e-3 * rng.uniform());
        rho[(size_t)ix + (size_t)n * (iy + (size_t)n * iz)] = v;
        sum += v;
      }
  const double scale = j.n4 / (sum * j.h * j.h * j.h);
  for (double& v : rho) v *= scale;
  return rho;
}

// potenimpini table (potenimpini.f90:31-59): npot = 100 n + 1 nodes on
// [0, 4 xmax] of an LJ pair potential capped inside r_cutoff = 3 A at its value
// there (Select_Pot semantics, He_statics/V_impur.f90:189-194)
std::vector<double> lj_table(const hedft_job& j, double* delinter, double* rmax) {
  const int npot = 100 * j.n + 1;
  *rmax = 4.0 * 0.5 * j.n * j.h;
  *delinter = *rmax / (npot - 1);
  const double rc = 3.0, sc6 = std::pow(j.sigma / rc, 6), umax = 4.0 * j.eps * (sc6 * sc6 - sc6);
  std::vector<double> t(npot);
  for (int i = 0; i < npot; ++i) {
    const double r = *delinter * i;
    if (r < rc) {
      t[i] = umax;
    } else {
      const double s6 = std::pow(j.sigma / r, 6);
      t[i] = 4.0 * j.eps * (s6 * s6 - s6);
    }
  }
  return t;
}

// V(r) of the capped LJ pair potential (the table above, evaluated exactly)
double lj_capped(const hedft_job& j, double r, double rc, double umax) {
  if (r < rc) return umax;
  const double s6 = std::pow(j.sigma / r, 6);
  return 4.0 * j.eps * (s6 * s6 - s6);
}

// Filon's rule for int_{x0}^{x0+2m h} f(x) sin(q x) dx (Abramowitz & Stegun
// 25.4.47), series weights for small theta = q h.  sin/cos of q x along the
// nodes follow by rotation, re-anchored every 64 nodes.
double filon_sin(const std::vector<double>& f, double x0, double h, double q) {
  const double t = q * h;
  double al, be, ga;
  if (t < 0.2) {
    const double t2 = t * t;
    al = t * t2 * (2.0 / 45.0 + t2 * (-2.0 / 315.0 + t2 * (2.0 / 4725.0)));
    be = 2.0 / 3.0 + t2 * (2.0 / 15.0 + t2 * (-4.0 / 105.0 + t2 * (2.0 / 567.0)));
    ga = 4.0 / 3.0 + t2 * (-2.0 / 15.0 + t2 * (1.0 / 210.0 + t2 * (-1.0 / 11340.0)));
  } else {
    const double sn = std::sin(t), cs = std::cos(t);
    al = 1.0 / t + sn * cs / (t * t) - 2.0 * sn * sn / (t * t * t);
    be = 2.0 * ((1.0 + cs * cs) / (t * t) - 2.0 * sn * cs / (t * t * t));
    ga = 4.0 * (sn / (t * t * t) - cs / (t * t));
  }
  const size_t n = f.size();  // 2m + 1 nodes
  const double ch = std::cos(t), sh = std::sin(t);
  double se = 0.0, so = 0.0, sk = 0.0, ck = 0.0;
  for (size_t k = 0; k < n; ++k) {
    if ((k & 63) == 0) {
      sk = std::sin(q * (x0 + h * k));
      ck = std::cos(q * (x0 + h * k));
    }
    if (k & 1) so += f[k] * sk;
    else se += f[k] * sk;
    const double s2 = sk * ch + ck * sh;
    ck = ck * ch - sk * sh;
    sk = s2;
  }
  const double xe = x0 + h * (n - 1);
  se -= 0.5 * (f[n - 1] * std::sin(q * xe) + f[0] * std::sin(q * x0));
  const double ends = f[0] * std::cos(q * x0) - f[n - 1] * std::cos(q * xe);
  return h * (al * ends + be * se + ga * so);
}

// uimp_k on the half spectrum (x fastest) of a cubic grid: the 3-D Fourier
// transform V^(p) = (4 pi/q) int_0^inf r V(r) sin(q r) dr, q = 2 pi |p|, of the
// capped LJ potential (potenimpini.f90:19-29 with FT_V_spline's definition):
// core r < 3 A in closed form, the tail by Filon's rule on [3, 50] A (step
// 0.005 A) and [50, 5000] A (step 0.5 A).  Values are computed once per
// distinct |p|^2 on the grid, in parallel host threads.
std::vector<double> uimp_k_lj(const hedft_job& j) {
  const int n = j.n, nxh = n / 2 + 1;
  const double rc = 3.0, sc6 = std::pow(j.sigma / rc, 6), umax = 4.0 * j.eps * (sc6 * sc6 - sc6);
  struct Seg {
    double x0, h;
    std::vector<double> f, f2;  // r V(r), r^2 V(r)
  };
  const double bounds[3] = {rc, 50.0, 5000.0}, steps[2] = {0.005, 0.5};
  Seg seg[2];
  for (int s = 0; s < 2; ++s) {
    int m = (int)std::ceil((bounds[s + 1] - bounds[s]) / steps[s]);
    m += m % 2;
    seg[s].x0 = bounds[s];
    seg[s].h = (bounds[s + 1] - bounds[s]) / m;
    seg[s].f.resize(m + 1);
    seg[s].f2.resize(m + 1);
    for (int k = 0; k <= m; ++k) {
      const double r = seg[s].x0 + seg[s].h * k;
      const double v = lj_capped(j, r, rc, umax);
      seg[s].f[k] = r * v;
      seg[s].f2[k] = r * r * v;
    }
  }
  const int half = n / 2, mmax = 3 * half * half;
  const double hp = 1.0 / (n * j.h), pi = 3.14159265358979323846;
  std::vector<char> used(mmax + 1, 0);
  for (int a = 0; a <= half; ++a)
    for (int b = 0; b <= half; ++b)
      for (int c = 0; c <= half; ++c) used[a * a + b * b + c * c] = 1;
  std::vector<double> vm(mmax + 1, 0.0);
  auto work = [&](int t0, int nt) {
    for (int m = t0; m <= mmax; m += nt) {
      if (!used[m]) continue;
      if (m == 0) {
        double tail = 0.0;
        for (const Seg& g : seg) {  // Simpson's rule
          double s = g.f2.front() + g.f2.back();
          for (size_t k = 1; k + 1 < g.f2.size(); ++k) s += ((k & 1) ? 4.0 : 2.0) * g.f2[k];
          tail += g.h / 3.0 * s;
        }
        vm[0] = 4.0 * pi * (umax * rc * rc * rc / 3.0 + tail);
        continue;
      }
      const double q = 2.0 * pi * std::sqrt((double)m) * hp;
      const double x = q * rc;
      const double core = umax * (std::sin(x) - x * std::cos(x)) / (q * q);
      double tail = 0.0;
      for (const Seg& g : seg) tail += filon_sin(g.f, g.x0, g.h, q);
      vm[m] = 4.0 * pi * (core + tail) / q;
    }
  };
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work, t, nt);
  work(0, nt);
  for (auto& x : th) x.join();
  auto fold = [&](int i) { return std::min(i, n - i); };  // |frequency index|
  std::vector<double> uk((size_t)nxh * n * n);
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < nxh; ++ix) {
        const int a = ix, b = fold(iy), c = fold(iz);
        uk[(size_t)ix + (size_t)nxh * (iy + (size_t)n * iz)] = vm[a * a + b * b + c * c];
      }
  return uk;
}

// the packet exp(-|r-r0|^2/(2 w^2) + i k.r), sum |psix|^2 dxyz = 1; k from the
// group velocity v0 (A/ps -> A K): v = 2 h2o2mimp k
std::vector<double2> wave_packet(const hedft_job& j, double h2o2mimp) {
  const int n = j.n;
  const double xmax = 0.5 * n * j.h, w = j.width;
  double k[3];
  for (int a = 0; a < 3; ++a) k[a] = j.v0[a] * kPstok / (2.0 * h2o2mimp);
  std::vector<double2> psi((size_t)n * n * n);
  double sum = 0.0;
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        const double x = -xmax + j.h * ix, y = -xmax + j.h * iy, z = -xmax + j.h * iz;
        const double dx = x - j.r0[0], dy = y - j.r0[1], dz = z - j.r0[2];
        const double amp = std::exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * w * w));
        const double ph = k[0] * x + k[1] * y + k[2] * z;
        psi[(size_t)ix + (size_t)n * (iy + (size_t)n * iz)] = make_double2(amp * std::cos(ph), amp * std::sin(ph));
        sum += amp * amp;
      }
  const double scale = 1.0 / std::sqrt(sum * j.h * j.h * j.h);
  for (double2& v : psi) {
    v.x *= scale;
    v.y *= scale;
  }
  return psi;
}

struct Active {
  int job;
  hedft_ctx* c;
  hipEvent_t ev;
  int steps_done;
  bool in_flight;
  size_t bytes;
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int admit(const hedft_job& j, int device, hedft_ctx** out) {
  if (j.n <= 0 || j.nsteps < 0 || j.nrk < 1 || j.h <= 0.0) return HEDFT_EINVAL;
  hedft_grid g{};
  g.nx = g.ny = g.nz = j.n;
  g.xmax = g.ymax = g.zmax = 0.5 * j.n * j.h;
  hedft_params p;
  hedft_default_params(&p);
  p.core_otc = 1;
  p.precision = j.precision;
  hedft_ctx* c = nullptr;
  int rc = hedft_create(&g, &p, HEDFT_MODE_TDDFT, device, &c);
  if (rc != HEDFT_OK) return rc;
  const std::vector<double> rho = droplet_density(j);
  if (rc == HEDFT_OK) rc = hedft_upload(c, HEDFT_DEN, rho.data(), rho.size() * sizeof(double));
  if (j.quantum) {
    if (!(j.width > 0.0) || !(j.mass_amu > 0.0)) rc = HEDFT_EINVAL;
    const size_t nh = (size_t)(j.n / 2 + 1) * j.n * j.n;
    std::vector<double> uk;
    if (rc == HEDFT_OK && !j.uimp_k) uk = uimp_k_lj(j);
    const double* ukp = j.uimp_k ? j.uimp_k : uk.data();
    if (rc == HEDFT_OK) rc = hedft_set_quantum_impurity(c, ukp, nh * sizeof(double), j.mass_amu, 0.0);
    if (rc == HEDFT_OK) {
      const std::vector<double2> px = wave_packet(j, c->qi[0].h2o2mimp);
      rc = hedft_upload(c, HEDFT_PSIX, px.data(), px.size() * sizeof(double2));
    }
  } else {
    double delinter = 0.0, rmax = 0.0;
    std::vector<double> tab;
    if (j.potion) {
      if (j.npot < 2 || !(j.delinter > 0.0)) rc = HEDFT_EINVAL;
      else {
        tab.assign(j.potion, j.potion + j.npot);
        delinter = j.delinter;
        rmax = j.delinter * (j.npot - 1);
      }
    } else {
      tab = lj_table(j, &delinter, &rmax);
    }
    const double mass = j.mass_amu * kMpU;
    const double r0[3] = {j.r0[0], j.r0[1], j.r0[2]};
    const double v0[3] = {j.v0[0] * kPstok, j.v0[1] * kPstok, j.v0[2] * kPstok};
    const double a0[3] = {0.0, 0.0, 0.0};
    if (rc == HEDFT_OK) rc = hedft_set_impurities(c, 1, (int32_t)tab.size(), delinter, rmax, tab.data(), &mass);
    if (rc == HEDFT_OK) rc = hedft_set_impurity_state(c, r0, v0, a0);
  }
  if (rc == HEDFT_OK) rc = hedft_tddft_init(c);
  for (int it = 1; it <= j.nrk && rc == HEDFT_OK; ++it)
    rc = hedft_tddft_iteration(c, j.dt, it, j.nrk, nullptr, nullptr, nullptr);
  if (rc != HEDFT_OK) {
    hedft_destroy(c);
    return rc;
  }
  *out = c;
  return HEDFT_OK;
}

void finish(const hedft_job& j, Active& a, hedft_job_result& r, double now) {
  hedft_energies e{};
  int rc = hedft_energy(a.c, &e);
  double rimp[3] = {0, 0, 0};
  if (rc == HEDFT_OK) {
    if (j.quantum) {
      double nrm = 0.0;
      rc = hedft_quantum_centroid(a.c, &nrm, rimp);
    } else {
      rc = hedft_get_impurity_state(a.c, rimp, nullptr, nullptr, nullptr, nullptr);
    }
  }
  r.status = rc;
  r.done_s = now;
  r.latency_s = now - j.arrival_s;
  r.etot = e.etot;
  r.etot4 = e.etot4;
  r.ekin4 = e.ekin4;
  r.norm = e.norm;
  for (int k = 0; k < 3; ++k) r.rimp[k] = rimp[k];
}

}  // namespace

extern "C" {

size_t hedft_job_bytes(const hedft_job* j) {
  if (!j) return 0;
  size_t b = hedft::ctx_footprint(j->n, HEDFT_MODE_TDDFT, j->precision);
  if (j->quantum) {  // psix, its PC history and Laplacian; potx4, upotx; uimp_k; 2 half spectra
    const size_t N = (size_t)j->n * j->n * j->n, Nh = (size_t)(j->n / 2 + 1) * j->n * j->n;
    const size_t fw = (j->precision == HEDFT_PREC_MIXED) ? 4 : 8;
    b += 9 * 16 * N + 2 * fw * N + fw * Nh + 2 * 2 * fw * Nh;
  }
  return b;
}

int hedft_serve(const hedft_job* jobs, int32_t njobs, const hedft_serve_opts* opts, hedft_job_result* results,
                hedft_serve_stats* stats) {
  if (!jobs || njobs < 0 || !opts || !results || opts->hbm_budget <= 0.0) return HEDFT_EINVAL;
  const int chunk = std::max(6, (opts->chunk + 5) / 6 * 6);
  std::vector<int> order(njobs);
  for (int i = 0; i < njobs; ++i) {
    order[i] = i;
    std::memset(&results[i], 0, sizeof(hedft_job_result));
    results[i].status = HEDFT_ESTATE;  // not run
    if ((double)hedft_job_bytes(&jobs[i]) > opts->hbm_budget) return HEDFT_EINVAL;  // could never be admitted
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
  if (hipSetDevice(opts->device) != hipSuccess) return HEDFT_EHIP;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<Active> active;
  size_t next = 0, used = 0, peak = 0;
  int max_conc = 0, finished = 0;
  long long total_steps = 0;
  int rc_all = HEDFT_OK;
  while ((next < order.size() || !active.empty()) && rc_all == HEDFT_OK) {
    double now = seconds_since(t0);
    bool progressed = false;
    // admission: arrival order, when the footprint fits the budget
    while (next < order.size()) {
      const hedft_job& j = jobs[order[next]];
      if (opts->realtime && j.arrival_s > now) break;
      const size_t b = hedft_job_bytes(&j);
      if ((double)(used + b) > opts->hbm_budget) break;
      if (opts->max_active > 0 && (int)active.size() >= opts->max_active) break;
      hedft_ctx* c = nullptr;
      const int rc = admit(j, opts->device, &c);
      hedft_job_result& r = results[order[next]];
      r.admit_s = seconds_since(t0);
      if (rc != HEDFT_OK) {
        r.status = rc;
        r.done_s = r.admit_s;
        ++next;
        continue;
      }
      Active a{order[next], c, nullptr, 0, false, b};
      if (hipEventCreateWithFlags(&a.ev, hipEventDisableTiming) != hipSuccess) {
        hedft_destroy(c);
        rc_all = HEDFT_EHIP;
        break;
      }
      active.push_back(a);
      used += b;
      peak = std::max(peak, used);
      max_conc = std::max(max_conc, (int)active.size());
      ++next;
      progressed = true;
    }
    // issue the next quantum of every idle job; retire finished ones
    for (size_t i = 0; i < active.size() && rc_all == HEDFT_OK;) {
      Active& a = active[i];
      const hedft_job& j = jobs[a.job];
      if (a.in_flight) {
        const hipError_t q = hipEventQuery(a.ev);
        if (q == hipErrorNotReady) {
          ++i;
          continue;
        }
        if (q != hipSuccess) {
          rc_all = HEDFT_EHIP;
          break;
        }
        a.in_flight = false;
        progressed = true;
      }
      if (a.steps_done >= j.nsteps) {
        finish(j, a, results[a.job], seconds_since(t0));
        total_steps += a.steps_done;
        ++finished;
        hipEventDestroy(a.ev);
        hedft_destroy(a.c);
        used -= a.bytes;
        active.erase(active.begin() + i);
        progressed = true;
        continue;
      }
      const int k = std::min(chunk, j.nsteps - a.steps_done);
      int rc = hedft::enqueue_pc_async(a.c, j.dt, k);
      if (rc == HEDFT_OK && hipEventRecord(a.ev, a.c->main_stream) != hipSuccess) rc = HEDFT_EHIP;
      if (rc != HEDFT_OK) {
        results[a.job].status = rc;
        hipStreamSynchronize(a.c->main_stream);
        hipEventDestroy(a.ev);
        hedft_destroy(a.c);
        used -= a.bytes;
        active.erase(active.begin() + i);
        continue;
      }
      a.steps_done += k;
      a.in_flight = true;
      progressed = true;
      ++i;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  for (Active& a : active) {  // only on a runtime error
    hipStreamSynchronize(a.c->main_stream);
    hipEventDestroy(a.ev);
    hedft_destroy(a.c);
  }
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->wall_s = seconds_since(t0);
    stats->steps_per_s = stats->wall_s > 0 ? total_steps / stats->wall_s : 0.0;
    std::vector<double> lat;
    for (int i = 0; i < njobs; ++i)
      if (results[i].status == HEDFT_OK && jobs[i].nsteps > 0)
        lat.push_back(results[i].latency_s * 1000.0 * 1000.0 / jobs[i].nsteps);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
      if (lat.empty()) return 0.0;
      const size_t k = (size_t)std::ceil(q * lat.size()) - 1;
      return lat[std::min(k, lat.size() - 1)];
    };
    stats->p50_ms_per_1000 = pct(0.50);
    stats->p99_ms_per_1000 = pct(0.99);
    stats->max_concurrent = max_conc;
    stats->peak_bytes = (double)peak;
    stats->completed = finished;
  }
  return rc_all;
}

}  // extern "C"
