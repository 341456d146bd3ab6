// serve.hip — continuous-batching request loop over independent TDDFT droplets.
//
// No counterpart in the reference, which runs one droplet per process.  Jobs
// (a droplet + classical impurity set-up and a number of PC iterations) arrive
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
  double delinter = 0.0, rmax = 0.0;
  const std::vector<double> tab = lj_table(j, &delinter, &rmax);
  const double mass = j.mass_amu * kMpU;
  const double r0[3] = {j.r0[0], j.r0[1], j.r0[2]};
  const double v0[3] = {j.v0[0] * kPstok, j.v0[1] * kPstok, j.v0[2] * kPstok};
  const double a0[3] = {0.0, 0.0, 0.0};
  if (rc == HEDFT_OK) rc = hedft_upload(c, HEDFT_DEN, rho.data(), rho.size() * sizeof(double));
  if (rc == HEDFT_OK) rc = hedft_set_impurities(c, 1, (int32_t)tab.size(), delinter, rmax, tab.data(), &mass);
  if (rc == HEDFT_OK) rc = hedft_set_impurity_state(c, r0, v0, a0);
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
  if (rc == HEDFT_OK) rc = hedft_get_impurity_state(a.c, rimp, nullptr, nullptr, nullptr, nullptr);
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
  return hedft::ctx_footprint(j->n, HEDFT_MODE_TDDFT, j->precision);
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
