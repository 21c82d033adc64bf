/*
 * cli_main.cpp — the command-line driver shared by the tri_matrix, n_array, grid
 * and tri_matrix_empty binaries (each links a different engine behind
 * cluster(), as the reference's Makefile:31-50 does).
 *
 * Reproduces the reference driver (main.cc) for users and scripts:
 *   flags    --ptmin/-p X, -r/-R X, --repeat/-repeat N, --sort/-s, --cartesian,
 *            --polar, --kt/-kt, --anti-kt/-antikt, --cambridge-aachen/-cam,
 *            --file/-f PATH, --combine/-combine N, --csv/-csv  (main.cc:129-257),
 *            with the same defaults (R 1.0, kt, ptmin 0, polar, 1 repetition,
 *            one event per clustering; main.cc:119-127) and the same messages and
 *            exit status 1 for a missing or unparsable argument or unknown flag;
 *   input    lines of "px py pz E"; lines starting with '#' and empty lines before
 *            an event are skipped; a '#' or empty line ends an event; any other
 *            unparsable line is an error (main.cc:34-85);
 *   timing   HIP events around host-to-device copy + cluster() + device-to-host
 *            copy, per repetition (main.cc:281-306);
 *   jets     the records with isJet and pT >= ptmin, sorted by decreasing pT, on the
 *            GPU (inclusive_jets of fjmi_dropin.h); only those come back, where the
 *            reference copies every record back and filters and sorts on the host
 *            (main.cc:309-327); the timed span ends after that copy;
 *   output   per repetition "found J jets [above P GeV] in T ms", then the jets
 *            printed as (rapidity, phi, pT) or
 *            (px, py, pz, E) with "%5d" and "%16.8f" columns (main.cc:87-116), then
 *            "clustered N particles into J jets above P GeV in T ms" with the mean
 *            and standard deviation of the repetitions, or the CSV line
 *            "N,J,mean[,sigma]" (main.cc:339-367).
 * Deliberate differences: the banner names the HIP device (and, where HIP does not
 * take a printf-buffer limit, states the 10 MiB requested); "--repeat 0" runs once
 * (the reference loops forever and divides by zero); a clustering error aborts
 * with the library's message.
 */
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "PseudoJet.h"
#include "cluster.h"
#include "fjmi_dropin.h"

namespace {

void hip_check(hipError_t e, const char *call, const char *file, int line) {
  if (e == hipSuccess) return;
  std::ostringstream out;
  out << "\n" << file << ", line " << line << ":\nhipCheck(" << call << ");\n"
      << hipGetErrorName(e) << ": " << hipGetErrorString(e) << "\n";
  std::cerr << out.str() << std::endl;
  std::abort();
}
#define HIP_CHECK(x) hip_check((x), #x, __FILE__, __LINE__)

void banner() {
  HIP_CHECK(hipSetDevice(0));
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, 0));
  std::cout << "Running on HIP device " << (prop.name[0] ? prop.name : prop.gcnArchName) << std::endl;
  std::cout << "  - maximum shared memory per block: " << prop.sharedMemPerBlock / 1024 << " kB" << std::endl;
  /* the reference asks for a 10 MiB device printf buffer and prints the size it got
   * (main.cc:28-31); HIP's device printf goes through host-call buffers, and where
   * the runtime does not take this limit the line states the size requested */
  size_t fifo = 10 * 1024 * 1024;
  if (hipDeviceSetLimit(hipLimitPrintfFifoSize, fifo) == hipSuccess) {
    size_t got = 0;
    if (hipDeviceGetLimit(&got, hipLimitPrintfFifoSize) == hipSuccess && got > 0) fifo = got;
  }
  (void)hipGetLastError(); /* an unsupported limit is not an error of this program */
  std::cout << "  - kernel printf buffer size:       " << fifo / 1024 << " kB" << std::endl;
}

/* Appends the next event's particles; returns whether one was found. */
bool next_event(std::istream &in, std::vector<PseudoJet> &parts) {
  in.clear();
  while (in.peek() == '#' || in.peek() == '\n') in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  int index = static_cast<int>(parts.size());
  bool any = false;
  std::string line;
  while (std::getline(in, line).good()) {
    std::istringstream fields(line);
    double px, py, pz, E;
    fields >> px >> py >> pz >> E;
    if (fields.fail()) {
      if (!line.empty() && line[0] != '#') throw std::runtime_error("Error while parsing particles:\n" + line);
      break;
    }
    parts.push_back({index++, false, px, py, pz, E});
    any = true;
  }
  return any;
}

/* Reads up to `count` events (0: all) into `parts`; returns how many. */
int next_events(std::istream &in, std::vector<PseudoJet> &parts, int count) {
  parts.clear();
  int n = 0;
  while ((count == 0 || n < count) && next_event(in, parts)) ++n;
  return n;
}

void print_jets(const std::vector<PseudoJet> &jets, bool cartesian) {
  std::cout << std::fixed << std::setprecision(8);
  int k = 0;
  for (const PseudoJet &j : jets) {
    std::cout << std::setw(5) << k++;
    if (cartesian) {
      std::cout << std::setw(16) << j.px << std::setw(16) << j.py << std::setw(16) << j.pz << std::setw(16) << j.E;
    } else {
      const double pt = std::hypot(j.px, j.py);
      double phi = std::atan2(j.py, j.px);
      while (phi > 2 * M_PI) phi -= 2 * M_PI;
      while (phi < 0.) phi += 2 * M_PI;
      const double m2 = std::max(0.0, (j.E + j.pz) * (j.E - j.pz) - pt * pt);
      const double epz = j.E + std::abs(j.pz);
      double y = 0.5 * std::log((pt * pt + m2) / (epz * epz));
      if (j.pz > 0) y = -y;
      std::cout << std::setw(16) << y << std::setw(16) << phi << std::setw(16) << pt;
    }
    std::cout << std::endl;
  }
  std::cout << std::endl;
}

struct Options {
  double ptmin = 0.0;
  double r = 1.0;
  Algorithm algo = Algorithm::Kt;
  bool sort = true;
  bool cartesian = false;
  long repetitions = 1;
  long combine = 1;
  std::string file;
  bool csv = false;
};

/* Returns -1 to continue, or the exit status. */
int parse(int argc, const char *argv[], Options &o) {
  auto value = [&](int &i) -> const char * {
    if (++i >= argc) {
      std::cerr << "Missing argument to option " << argv[i - 1] << std::endl;
      return nullptr;
    }
    return argv[i];
  };
  auto bad = [&](int i) {
    std::cerr << "Error while parsing argument to option " << argv[i - 1] << std::endl;
    return 1;
  };
  auto is = [&](int i, const char *a, const char *b) {
    return std::strcmp(argv[i], a) == 0 || (b && std::strcmp(argv[i], b) == 0);
  };
  for (int i = 1; i < argc; ++i) {
    if (is(i, "--ptmin", "-p") || is(i, "-r", "-R")) {
      const bool pt = is(i, "--ptmin", "-p");
      const char *v = value(i);
      if (!v) return 1;
      char *stop;
      const double x = std::strtod(v, &stop);
      if (stop == v || !(x >= 0.)) return bad(i);
      (pt ? o.ptmin : o.r) = x;
    } else if (is(i, "--repeat", "-repeat") || is(i, "--combine", "-combine")) {
      const bool rep = is(i, "--repeat", "-repeat");
      const char *v = value(i);
      if (!v) return 1;
      char *stop;
      const long x = std::strtol(v, &stop, 0);
      if (stop == v || x < 0) return bad(i);
      (rep ? o.repetitions : o.combine) = x;
    } else if (is(i, "--sort", "-s")) {
      o.sort = true;
    } else if (is(i, "--cartesian", nullptr)) {
      o.cartesian = true;
    } else if (is(i, "--polar", nullptr)) {
      o.cartesian = false;
    } else if (is(i, "--kt", "-kt")) {
      o.algo = Algorithm::Kt;
    } else if (is(i, "--anti-kt", "-antikt")) {
      o.algo = Algorithm::AntiKt;
    } else if (is(i, "--cambridge-aachen", "-cam")) {
      o.algo = Algorithm::CambridgeAachen;
    } else if (is(i, "--file", "-f")) {
      const char *v = value(i);
      if (!v) return 1;
      o.file = v;
    } else if (is(i, "--csv", "-csv")) {
      o.csv = true;
    } else {
      std::cerr << "Unrecognized option " << argv[i] << std::endl;
      return 1;
    }
  }
  return -1;
}

}  // namespace

int main(int argc, const char *argv[]) {
  Options o;
  const int rc = parse(argc, argv, o);
  if (rc >= 0) return rc;
  const long reps = o.repetitions > 0 ? o.repetitions : 1;

  banner();
  std::ifstream file(o.file, std::ios_base::in);
  std::istream &in = o.file.empty() ? std::cin : file;

  std::vector<PseudoJet> parts, jets;
  int events;
  while ((events = next_events(in, parts, static_cast<int>(o.combine))) != 0) {
    if (!o.csv) {
      std::cout << "found " << parts.size() << " particles";
      if (o.combine != 1) std::cout << " in " << events << (events == 1 ? " event" : " events");
      std::cout << std::endl;
    }
    const size_t bytes = sizeof(PseudoJet) * parts.size();
    PseudoJet *dev = nullptr, *dev_jets = nullptr;
    int *dev_count = nullptr;
    HIP_CHECK(hipMalloc(&dev, bytes));
    HIP_CHECK(hipMalloc(&dev_jets, bytes > 0 ? bytes : sizeof(PseudoJet)));
    HIP_CHECK(hipMalloc(&dev_count, sizeof(int)));
    hipEvent_t start, stop;
    HIP_CHECK(hipEventCreate(&start));
    HIP_CHECK(hipEventCreate(&stop));
    double sum = 0., sum2 = 0.;
    for (long step = 0; step < reps; ++step) {
      HIP_CHECK(hipEventRecord(start, nullptr));
      HIP_CHECK(hipMemcpy(dev, parts.data(), bytes, hipMemcpyHostToDevice));
      cluster(dev, static_cast<int>(parts.size()), o.algo, o.r);
      /* the jets above ptmin, sorted by decreasing pT, on the device; only they come back */
      inclusive_jets(dev, static_cast<int>(parts.size()), o.ptmin, dev_jets, dev_count);
      int count = 0;
      HIP_CHECK(hipMemcpy(&count, dev_count, sizeof count, hipMemcpyDeviceToHost));
      jets.resize(static_cast<size_t>(count));
      if (count > 0) HIP_CHECK(hipMemcpy(jets.data(), dev_jets, sizeof(PseudoJet) * count, hipMemcpyDeviceToHost));
      HIP_CHECK(hipEventRecord(stop, nullptr));
      HIP_CHECK(hipEventSynchronize(stop));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
      sum += ms;
      sum2 += ms * ms;
      if (!o.csv) {
        if (o.ptmin > 0.)
          std::cout << "found " << jets.size() << " jets above " << o.ptmin << " GeV in " << ms << " ms" << std::endl;
        else
          std::cout << "found " << jets.size() << " jets in " << ms << " ms" << std::endl;
      }
    }
    HIP_CHECK(hipEventDestroy(start));
    HIP_CHECK(hipEventDestroy(stop));
    HIP_CHECK(hipFree(dev));
    HIP_CHECK(hipFree(dev_jets));
    HIP_CHECK(hipFree(dev_count));

    if (!o.csv) print_jets(jets, o.cartesian);
    std::cout << std::defaultfloat;
    if (!o.csv)
      std::cout << "clustered " << parts.size() << " particles into " << jets.size() << " jets above " << o.ptmin
                << " GeV";
    else
      std::cout << parts.size() << ',' << jets.size() << ',';
    std::cout << std::fixed;
    const double mean = sum / reps;
    std::streamsize saved;
    if (reps > 1) {
      const double sigma = std::sqrt((sum2 - sum * sum / reps) / (reps - 1));
      saved = std::cout.precision(std::max(static_cast<int>(-std::log10(sigma / 2.)) + 1, 0));
      if (!o.csv)
        std::cout << " in " << mean << " +/- " << sigma << " ms" << std::endl;
      else
        std::cout << mean << ',' << sigma << std::endl;
    } else {
      saved = std::cout.precision(1);
      if (!o.csv)
        std::cout << " in " << mean << " ms" << std::endl;
      else
        std::cout << mean << std::endl;
    }
    std::cout.precision(saved);
    std::cout << std::defaultfloat;
  }
  return 0;
}
