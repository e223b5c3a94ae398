// Fortran-callable named timers and the inert analytic-integral symbols.
//
// Timers keep the reference's observable behaviour (g2g/timer.cpp:125-296):
// with an odd timer level, g2g_timer_stop_/pause_ print "TIMER [name]: t";
// with a level above 1 the sum timers nest under the currently running one
// and g2g_timer_summary_ prints the tree with each child's share of its
// parent.  The device is synchronised before every reading, as the
// reference's Timer::sync does, so GPU work is attributed correctly.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <map>
#include <set>
#include <string>

#include "../../include/g2g.h"

namespace xcq {

namespace {

struct Clock {
  double total = 0.0, started = 0.0;
  bool running = false;
  void start() {
    started = now();
    running = true;
  }
  void stop() {  // stop accumulates, like pause, and ends the interval
    if (running) total += now() - started;
    running = false;
  }
  static double now() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
  }
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

std::string name_of(const char* s, unsigned len) {
  std::string n(s, len);
  while (!n.empty() && n.back() == ' ') n.pop_back();
  return n;
}

void print_tree(const std::string& indent, const std::string& name, double parent_total,
                const std::string& parent) {
  const double t = g_all[name].total;
  std::printf("%s%-35s%12.6fs (%6.2f%% of %s)\n", indent.c_str(), name.c_str(), t,
              parent_total > 0 ? 100.0 * t / parent_total : 0.0, parent.c_str());
  for (const auto& c : g_children[name]) print_tree(indent + "  ", c, t, name);
}

}  // namespace

void timers_configure(unsigned level) {
  g_sum = level > 1;
  g_single = (level % 2) == 1;
}

}  // namespace xcq

using namespace xcq;

extern "C" {

void g2g_timer_start_(const char* name, unsigned len) {
  if (!g_single) return;
  sync_device();
  g_single_timers[name_of(name, len)].start();
}

void g2g_timer_stop_(const char* name, unsigned len) {
  if (!g_single) return;
  sync_device();
  const std::string n = name_of(name, len);
  if (!g_single_timers.count(n)) std::printf("no such timer (%s)\n", n.c_str());
  Clock& c = g_single_timers[n];
  c.stop();
  std::printf("TIMER [%s]: %.6fs\n", n.c_str(), c.total);
  c.total = 0.0;
}

void g2g_timer_pause_(const char* name, unsigned len) {
  if (!g_single) return;
  sync_device();
  const std::string n = name_of(name, len);
  if (!g_single_timers.count(n)) std::printf("no such timer (%s)\n", n.c_str());
  Clock& c = g_single_timers[n];
  c.stop();
  std::printf("TIMER [%s]: %.6fs(so far)\n", n.c_str(), c.total);
}

void g2g_timer_sum_start_(const char* name, unsigned len) {
  if (!g_sum) return;
  const std::string n = name_of(name, len);
  g_children[n];
  if (g_current.empty()) {
    g_top.insert(n);
  } else if (!g_children[g_current].count(n) && !g_parent.count(n)) {
    g_children[g_current].insert(n);
    g_parent[n] = g_current;
  }
  sync_device();
  g_all[n].start();
  g_current = n;
}

static void sum_end(const char* name, unsigned len) {
  const std::string n = name_of(name, len);
  sync_device();
  if (g_current != n) {
    std::printf("Error: not the current timer: (%s,%s)\n", n.c_str(), g_current.c_str());
    return;
  }
  g_all[n].stop();
  auto it = g_parent.find(n);
  g_current = (it == g_parent.end()) ? std::string() : it->second;
}

void g2g_timer_sum_stop_(const char* name, unsigned len) {
  if (g_sum) sum_end(name, len);
}

void g2g_timer_sum_pause_(const char* name, unsigned len) {
  if (g_sum) sum_end(name, len);
}

void g2g_timer_clear_(void) {
  if (!g_sum) return;
  g_all.clear();
  g_children.clear();
  g_parent.clear();
  g_top.clear();
  g_current.clear();
}

void g2g_timer_summary_(void) {
  if (!g_sum) return;
  const double total = g_all.count("Total") ? g_all["Total"].total : 0.0;
  std::printf("-------------------------------------------------------------------------------------------\n");
  std::printf("                   LIO TIMING INFORMATION:\n");
  std::printf("-------------------------------------------------------------------------------------------\n\n");
  std::printf("Total time: %gs\n", total);
  for (const auto& t : g_top) print_tree("", t, total, "Total");
  std::printf("-------------------------------------------------------------------------------------------\n");
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
void aint_qmmm_fock_(double*, double*) {}
void aint_coulomb_forces_(double*) {}
void aint_coulomb_fock_(double*) {}

}  // extern "C"
