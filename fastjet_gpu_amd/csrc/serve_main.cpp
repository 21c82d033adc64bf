/*
 * serve_main.cpp — fjmi_serve: open-loop load generator for the serving path
 * (fjmi_server_*, include/fjmi.h).  Not in the reference, which clusters one
 * event per call (main.cc:269-368).
 *
 * Requests arrive as a Poisson process of rate lambda: request i arrives at
 * sum_{k<=i} gap_k / lambda with unit-mean exponential gaps, and takes its
 * algorithm and radius uniformly from {anti-kt 0.4, anti-kt 0.8, C/A 0.8, kt 0.4}
 * (fjmi_gen_serving_attrs).  Its particles are those of generator workload 5
 * (multiplicity lognormal with median 1500 and sigma 0.6, capped at 20000),
 * event i mod --unique, generated before the clock starts.  Arrived requests
 * wait in a host FIFO until the server admits them (memory budget and event
 * slots); completed events are released as soon as they are seen.
 *
 * Latency of a request = the time its completion is seen by the host loop minus
 * its arrival time (so it includes queueing for admission and on the GPU).
 * Throughput = requests / (last completion - first arrival); beside it, the rate
 * between the 10th and 90th percentile completions (without the ramp-up and
 * the final drain, whose largest events finish alone).
 *
 *   fjmi_serve [--rates 1e3,1e4,1e5,1e6] [--events N] [--seconds S] [--unique U]
 *              [--device-gib G] [--host-gib H] [--slots E] [--variant V]
 *              [--replay FILE]
 * One JSON line per rate.  --events 0 (default) uses min(200000, rate * S)
 * requests with S = --seconds (20).  --variant: the tiled kernel variant
 * (default 1, 384 tiles and 10 workgroups per CU: the higher served rate near
 * saturation; variant 0 gives up to 2 % lower p99 at light load).  --replay:
 * the events of a binary event file (fjmi_events_*) replace the generated
 * ones, request i taking event i mod E and, where the file holds them, that
 * event's exponent and radius.
 */
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "fjmi.h"
#include "fjmi_gen.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

[[noreturn]] void die(const char *what) {
  std::fprintf(stderr, "fjmi_serve: %s: %s\n", what, fjmi_last_error());
  std::exit(1);
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)pos;
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

struct Request {
  int event;
  int p;
  double R;
  double arrival;
};

}  // namespace

int main(int argc, char **argv) {
  std::vector<double> rates = {1e3, 1e4, 1e5, 1e6};
  long events_arg = 0;
  double seconds = 20.0;
  int unique = 2048;
  double device_gib = 16.0, host_gib = 8.0;
  int slots = 0, variant = 1;
  std::string replay;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "Missing argument to option %s\n", a.c_str());
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "--rates") {
      rates.clear();
      std::string s = next();
      size_t pos = 0;
      while (pos < s.size()) {
        size_t c = s.find(',', pos);
        if (c == std::string::npos) c = s.size();
        rates.push_back(std::atof(s.substr(pos, c - pos).c_str()));
        pos = c + 1;
      }
    } else if (a == "--events") {
      events_arg = std::atol(next());
    } else if (a == "--seconds") {
      seconds = std::atof(next());
    } else if (a == "--unique") {
      unique = std::atoi(next());
    } else if (a == "--device-gib") {
      device_gib = std::atof(next());
    } else if (a == "--host-gib") {
      host_gib = std::atof(next());
    } else if (a == "--slots") {
      slots = std::atoi(next());
    } else if (a == "--variant") {
      variant = std::atoi(next());
    } else if (a == "--replay") {
      replay = next();
    } else {
      std::fprintf(stderr, "Unrecognized option %s\n", a.c_str());
      return 1;
    }
  }
  if (unique < 1 || rates.empty()) {
    std::fprintf(stderr, "need --unique >= 1 and at least one rate\n");
    return 1;
  }
  for (double r : rates)
    if (!(r > 0.0)) {
      std::fprintf(stderr, "rates must be positive\n");
      return 1;
    }

  /* the distinct events, generated or read before any clock starts */
  std::vector<std::vector<double>> pool;
  std::vector<int32_t> file_p;
  std::vector<double> file_R;
  if (!replay.empty()) {
    int ne = 0, has_p = 0, has_R = 0;
    int64_t np = 0;
    if (fjmi_events_info(replay.c_str(), &ne, &np, &has_p, &has_R)) die("fjmi_events_info");
    if (ne < 1) {
      std::fprintf(stderr, "fjmi_serve: %s holds no events\n", replay.c_str());
      return 1;
    }
    std::vector<int64_t> off((size_t)ne + 1);
    std::vector<double> p4(4 * (size_t)std::max<int64_t>(np, 1));
    if (has_p) file_p.resize(ne);
    if (has_R) file_R.resize(ne);
    if (fjmi_events_read(replay.c_str(), off.data(), p4.data(), has_p ? file_p.data() : nullptr,
                         has_R ? file_R.data() : nullptr))
      die("fjmi_events_read");
    for (int e = 0; e < ne; ++e)
      if (off[e + 1] - off[e] > 2000000000 / 4) {
        std::fprintf(stderr, "fjmi_serve: event %d of %s is too large\n", e, replay.c_str());
        return 1;
      }
    pool.resize(ne);
    for (int e = 0; e < ne; ++e) pool[e].assign(p4.begin() + 4 * off[e], p4.begin() + 4 * off[e + 1]);
    unique = ne;
  } else {
    pool.resize(unique);
    for (int e = 0; e < unique; ++e) {
      const int n = fjmi_gen_event_size(5, e);
      pool[e].resize(4 * (size_t)std::max(n, 1));
      if (fjmi_gen_event(5, e, pool[e].data(), n) != n) die("generator");
      pool[e].resize(4 * (size_t)n);
    }
  }
  double mean_n = 0.0;
  for (const auto &ev : pool) mean_n += (double)(ev.size() / 4);
  mean_n /= unique;

  fjmi_context *ctx = nullptr;
  if (fjmi_create(0, &ctx)) die("fjmi_create");
  fjmi_server_config cfg{};
  cfg.device_budget = (int64_t)(device_gib * (double)(1ll << 30));
  cfg.host_budget = (int64_t)(host_gib * (double)(1ll << 30));
  cfg.max_events = slots;
  cfg.variant = variant;
  const auto t_create = Clock::now();
  fjmi_server *srv = nullptr;
  if (fjmi_server_create(ctx, &cfg, &srv)) die("fjmi_server_create");
  const double create_s = seconds_since(t_create);
  fjmi_server_info info0{};
  fjmi_server_info_get(srv, &info0);

  std::vector<fjmi_server_result> res(4096);
  std::vector<double> done_at; /* completion times of the current rate's requests */
  /* no completion for this long while requests are outstanding ends the run */
  const double stall_limit_s = 30.0;
  double last_progress = 0.0;
  auto check_stall = [&](Clock::time_point t0, long outstanding) {
    const double now = seconds_since(t0);
    if (outstanding > 0 && now - last_progress > stall_limit_s) {
      std::fprintf(stderr, "fjmi_serve: no completion for %.0f s with %ld requests outstanding\n", stall_limit_s,
                   outstanding);
      std::exit(2);
    }
  };
  auto drain = [&](std::vector<double> *lat, const std::vector<Request> *reqs, Clock::time_point t0,
                   double *last_done, long *done) {
    int k = 0;
    if (fjmi_server_poll(srv, res.data(), (int)res.size(), &k)) die("fjmi_server_poll");
    const double now = seconds_since(t0);
    for (int i = 0; i < k; ++i) {
      if (res[i].njets < 0) {
        std::fprintf(stderr, "fjmi_serve: request %llu failed on the GPU\n", (unsigned long long)res[i].tag);
        std::exit(1);
      }
      if (lat) {
        (*lat)[res[i].tag] = now - (*reqs)[res[i].tag].arrival;
        done_at.push_back(now);
      }
      if (fjmi_server_release(srv, res[i].handle)) die("fjmi_server_release");
    }
    if (k) {
      *last_done = now;
      last_progress = now;
    }
    *done += k;
    return k;
  };

  /* warm-up: kernels loaded, first launch done (not timed) */
  {
    const auto t0 = Clock::now();
    double last = 0.0;
    long done = 0;
    last_progress = 0.0;
    for (int e = 0; e < 64; ++e) {
      const std::vector<double> &ev = pool[e % unique];
      int rc;
      while ((rc = fjmi_server_submit(srv, ev.data(), (int)(ev.size() / 4), -1, 0.4, (uint64_t)e)) == FJMI_EAGAIN) {
        drain(nullptr, nullptr, t0, &last, &done);
        check_stall(t0, e - done);
      }
      if (rc) die("fjmi_server_submit");
    }
    while (done < 64) {
      drain(nullptr, nullptr, t0, &last, &done);
      check_stall(t0, 64 - done);
    }
  }

  for (double rate : rates) {
    const long n_req = events_arg > 0 ? events_arg : std::min<long>(200000, std::max<long>(1000, (long)(rate * seconds)));
    std::vector<Request> reqs(n_req);
    double t = 0.0;
    for (long i = 0; i < n_req; ++i) {
      int p;
      double R, gap;
      fjmi_gen_serving_attrs(i, &p, &R, &gap);
      t += gap / rate;
      const int ev = (int)(i % unique);
      if (!file_p.empty()) p = file_p[ev];
      if (!file_R.empty()) R = file_R[ev];
      reqs[i] = Request{ev, p, R, t};
    }
    std::vector<double> lat(n_req, 0.0), admit_wait(n_req, 0.0);
    done_at.clear();
    done_at.reserve(n_req);
    std::deque<long> fifo;
    long next = 0, done = 0, eagain = 0, max_fifo = 0;
    double last_done = 0.0;
    fjmi_server_info before{};
    fjmi_server_info_get(srv, &before);
    const auto t0 = Clock::now();
    last_progress = 0.0;
    double last_report = 0.0;
    while (done < n_req) {
      check_stall(t0, next - done);
      if (seconds_since(t0) - last_report > 10.0) { /* progress line for long runs */
        last_report = seconds_since(t0);
        std::fprintf(stderr, "[rate %g] %.0f s: %ld of %ld requests done\n", rate, last_report, done, n_req);
      }
      const double now = seconds_since(t0);
      while (next < n_req && reqs[next].arrival <= now) fifo.push_back(next++);
      max_fifo = std::max<long>(max_fifo, (long)fifo.size());
      while (!fifo.empty()) {
        const Request &q = reqs[fifo.front()];
        const std::vector<double> &ev = pool[q.event];
        const int rc = fjmi_server_submit(srv, ev.data(), (int)(ev.size() / 4), q.p, q.R, (uint64_t)fifo.front());
        if (rc == FJMI_EAGAIN) {
          ++eagain;
          break;
        }
        if (rc) die("fjmi_server_submit");
        admit_wait[fifo.front()] = seconds_since(t0) - q.arrival;
        fifo.pop_front();
      }
      drain(&lat, &reqs, t0, &last_done, &done);
    }
    fjmi_server_info after{};
    fjmi_server_info_get(srv, &after);
    const double span = last_done - reqs[0].arrival;
    /* the served rate between the 10th and the 90th percentile completion: the
     * server's rate once started and before the last events drain (an overloaded
     * run ends with a backlog whose largest events finish alone) */
    std::sort(done_at.begin(), done_at.end());
    const size_t i10 = done_at.size() / 10, i90 = done_at.size() * 9 / 10;
    const double mid = (i90 > i10 && done_at[i90] > done_at[i10])
                           ? (double)(i90 - i10) / (done_at[i90] - done_at[i10]) : 0.0;
    std::printf(
        "{\"rate\": %.6g, \"requests\": %ld, \"events_per_s\": %.1f, \"mid_events_per_s\": %.1f, \"offered_per_s\": %.1f, "
        "\"latency_ms\": {\"p50\": %.3f, \"p90\": %.3f, \"p99\": %.3f, \"p999\": %.3f, \"max\": %.3f}, "
        "\"admission_wait_ms\": {\"p50\": %.3f, \"p99\": %.3f}, \"max_host_queue\": %ld, \"refused_submits\": %ld, "
        "\"peak_particles_in_flight\": %lld, \"particle_capacity\": %lld, \"event_slots\": %d, "
        "\"device_bytes\": %lld, \"host_bytes\": %lld, \"launches\": %lld, \"grid\": %d, "
        "\"mean_n_unique\": %.1f, \"unique_events\": %d, \"server_create_s\": %.3f, \"gpu_peak_busy\": %lld}\n",
        rate, n_req, (double)n_req / span, mid, (double)n_req / reqs[n_req - 1].arrival, 1e3 * percentile(lat, 0.50),
        1e3 * percentile(lat, 0.90), 1e3 * percentile(lat, 0.99), 1e3 * percentile(lat, 0.999),
        1e3 * percentile(lat, 1.0), 1e3 * percentile(admit_wait, 0.50), 1e3 * percentile(admit_wait, 0.99), max_fifo,
        eagain, (long long)after.peak_particles_in_flight, (long long)after.particle_capacity, after.event_capacity,
        (long long)after.device_bytes, (long long)after.host_bytes, (long long)(after.launches - before.launches),
        after.grid, mean_n, unique, create_s, (long long)after.gpu_peak_busy);
    std::fflush(stdout);
  }
  if (fjmi_server_destroy(srv)) die("fjmi_server_destroy");
  fjmi_destroy(ctx);
  return 0;
}
