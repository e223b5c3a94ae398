// Fortran-callable named timers and the inert analytic-integral symbols.
//
// Timers keep the reference's observable behaviour (g2g/timer.cpp:37-112,
// 125-296):
//  * timer level odd: g2g_timer_start_ starts a named timer, g2g_timer_stop_
//    sets it to the interval since that start and prints
//    "TIMER [name]: <t>", g2g_timer_pause_ adds the interval to what the timer
//    holds and prints "TIMER [name]: <t>(so far)", where <t> is "<s>s. <us>us."
//    or "<us>us." below one second;
//  * timer level above 1: the sum timers nest under the one currently
//    running (a name is registered under the first parent it runs in), stop
//    sets and pause accumulates as above, and g2g_timer_summary_ prints the
//    tree, sorted by name, with each timer's share of its parent ("Total"
//    for the top level);
//  * the levels only switch timers on, as in g2g_set_options_
//    (g2g/init.cpp:427-432).
// Names keep the exact length Fortran passes.  The device is synchronised
// before every reading, as the reference's Timer::sync does, so GPU work is
// attributed correctly.  Output is flushed at once so it interleaves with
// the caller's own writes.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <map>
#include <set>
#include <string>

#include "../../include/g2g.h"

namespace xcq {

namespace {

// Nanosecond clock with the reference Timer's start / stop / pause meaning.
struct Clock {
  long long res = 0, t0 = 0;
  bool started = false;
  static long long now() {
    using namespace std::chrono;
    return duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count();
  }
  void start() {
    started = true;
    t0 = now();
  }
  void stop() {  // the interval since the last start
    res = now() - t0;
    started = false;
  }
  void pause() {  // adds the interval since the last start
    res += now() - t0;
    t0 = now();
  }
  unsigned long long sec() const { return (unsigned long long)(res / 1000000000LL); }
  unsigned long long usec() const { return (unsigned long long)((res % 1000000000LL) / 1000); }
  float seconds() const { return (float)sec() + (float)usec() / 1000000.0f; }
};

bool g_single = false, g_sum = false;
std::map<std::string, Clock> g_single_timers;
std::map<std::string, Clock> g_all;
std::map<std::string, std::set<std::string>> g_children;
std::map<std::string, std::string> g_parent;
std::set<std::string> g_top;
std::string g_current;

void sync_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) == hipSuccess && n > 0) (void)hipDeviceSynchronize();
}

void print_time(const Clock& c) {
  if (c.sec() != 0)
    std::printf("%llus. %lluus.", c.sec(), c.usec());
  else
    std::printf("%lluus.", c.usec());
}

void print_tree(const std::string& indent, const std::string& name, float parent_total,
                const std::string& parent) {
  const float t = g_all[name].seconds();
  std::printf("%s%-35s%12.6fs (%6.2f%% of %s)\n", indent.c_str(), name.c_str(), t,
              100.0f * t / parent_total, parent.c_str());
  for (const auto& c : g_children[name]) print_tree(indent + "  ", c, t, name);
}

}  // namespace

void timers_configure(unsigned level) {
  if (level > 1) g_sum = true;
  if (level % 2 == 1) g_single = true;
}

}  // namespace xcq

using namespace xcq;

extern "C" {

void g2g_timer_start_(const char* name, unsigned len) {
  if (!g_single) return;
  const std::string n(name, len);
  sync_device();
  g_single_timers[n].start();
}

static void single_end(const char* name, unsigned len, bool pause) {
  const std::string n(name, len);
  sync_device();
  if (!g_single_timers.count(n)) std::printf("g2g timer %s was never started\n", n.c_str());
  Clock& c = g_single_timers[n];
  if (pause)
    c.pause();
  else
    c.stop();
  std::printf("TIMER [%s]: ", n.c_str());
  print_time(c);
  std::printf(pause ? "(so far)\n" : "\n");
  std::fflush(stdout);
}

void g2g_timer_stop_(const char* name, unsigned len) {
  if (g_single) single_end(name, len, false);
}

void g2g_timer_pause_(const char* name, unsigned len) {
  if (g_single) single_end(name, len, true);
}

void g2g_timer_sum_start_(const char* name, unsigned len) {
  if (!g_sum) return;
  const std::string n(name, len);
  g_children[n];
  if (g_current.empty()) {
    g_top.insert(n);
  } else if (!g_children[g_current].count(n)) {
    g_children[g_current].insert(n);
    g_all[n] = Clock();  // registered anew under this parent
    g_parent[n] = g_current;
  }
  sync_device();
  g_all[n].start();
  g_current = n;
}

static void sum_end(const char* name, unsigned len, bool pause) {
  const std::string n(name, len);
  sync_device();
  if (g_current != n) {
    if (pause)
      std::printf("Error: not the current timer: (%s)\n", n.c_str());
    else
      std::printf("Error: not the current timer: (%s,%s)\n", n.c_str(), g_current.c_str());
    std::fflush(stdout);
    return;
  }
  if (pause)
    g_all[n].pause();
  else
    g_all[n].stop();
  auto it = g_parent.find(n);
  g_current = (it == g_parent.end()) ? std::string() : it->second;
}

void g2g_timer_sum_stop_(const char* name, unsigned len) {
  if (g_sum) sum_end(name, len, false);
}

void g2g_timer_sum_pause_(const char* name, unsigned len) {
  if (g_sum) sum_end(name, len, true);
}

void g2g_timer_clear_(void) {
  if (!g_sum) return;
  g_all.clear();
  g_children.clear();
  g_parent.clear();
  g_top.clear();
}

void g2g_timer_summary_(void) {
  if (!g_sum) return;
  // a summary without a "Total" timer reports a zero total (the reference
  // dereferences a missing entry there)
  const float total = g_all.count("Total") ? g_all["Total"].seconds() : 0.0f;
  const char* rule =
      "-------------------------------------------------------------------------------------------";
  std::printf("%s\n                   LIO TIMING INFORMATION:\n%s\n\n", rule, rule);
  std::printf("Total time: %gs\n", total);
  for (const auto& t : g_top) print_tree("", t, total, "Total");
  std::printf("%s\n", rule);
  std::fflush(stdout);
}

// Analytic integrals are not part of this library: report GPU level 0 so
// lioamber uses its CPU integrals (g2g/analytic_integral/aint_init.cpp:342-345).
void aint_query_gpu_level_(int* gpu_level) { *gpu_level = 0; }
void aint_parameter_init_(const unsigned*, unsigned*, const unsigned*, double*, double*, unsigned*,
                          double*, double*, double*, double*, double*, double*, unsigned*,
                          const int*) {}
void aint_deinit_(void) {}
void aint_new_step_(void) {}
void aint_qmmm_init_(const unsigned*, double*, double*) {}
void aint_coulomb_init_(void) {}
void aint_qmmm_forces_(double*, double*) {}
// The Fock terms are zeroed like the reference's (aint_init.cpp:267-269, 326-327)
// and nothing is added: with GPU level 0 lioamber computes them on the CPU.
void aint_qmmm_fock_(double* Es, double* Ens) {
  *Es = 0.0;
  *Ens = 0.0;
}
void aint_coulomb_forces_(double*) {}
void aint_coulomb_fock_(double* Es) { *Es = 0.0; }

}  // extern "C"
