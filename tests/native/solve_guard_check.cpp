// Host check of the pressure solves' residual watch (pyro_amd/csrc/solve_guard.hpp):
// residual sequences driven through the control flow of the solve loops in
// poisson.hip (initial residual, then `while rel > tol && it < max_iters`).
// Prints one line per case and exits nonzero on any mismatch.
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "solve_guard.hpp"

using pyro::ResidWatch;
using pyro::Watch;

struct Outcome {
  int iterations;
  Watch status;
};

// seq[0] is the initial residual, seq[k] the residual after iteration k
static Outcome run(const std::vector<double> &seq, double tol, int max_iters) {
  ResidWatch rw;
  int it = 0;
  double rel = seq[0];
  Watch w = rw.step(rel, 0);
  while (w == Watch::Go && rel > tol && it < max_iters) {
    it++;
    rel = seq[(size_t)it < seq.size() ? it : seq.size() - 1];
    w = rw.step(rel, it);
  }
  return {it, w};
}

static int bad = 0;
static void expect(const char *name, Outcome o, int it, Watch w) {
  const bool ok = o.iterations == it && o.status == w;
  std::printf("%-34s iterations %3d status %d  %s\n", name, o.iterations, (int)o.status, ok ? "ok" : "WRONG");
  if (!ok) bad++;
}

int main() {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<double> conv;
  for (int k = 0; k < 40; k++) conv.push_back(std::pow(0.3, k));
  expect("converges to 1e-6", run(conv, 1e-6, 100), 12, Watch::Go);
  expect("iteration cap", run(conv, 1e-30, 5), 5, Watch::Go);
  expect("NaN after 3 iterations", run({1.0, 0.5, 0.4, nan, 0.1}, 1e-6, 100), 3, Watch::NonFinite);
  expect("NaN initial residual", run({nan, 0.5}, 1e-6, 100), 0, Watch::NonFinite);
  expect("Inf after 2 iterations", run({1.0, 0.5, inf}, 1e-6, 100), 2, Watch::NonFinite);
  // growth by 1.1 per cycle from 1e-3 (the plain V-cycle at 256^3): stops at
  // the first residual above 1e3 x 1e-3
  std::vector<double> grow;
  for (int k = 0; k < 200; k++) grow.push_back(1e-3 * std::pow(1.1, k));
  int first = 0;
  while (!(grow[first] > 1.0)) first++;
  expect("geometric growth 1.1x", run(grow, 1e-6, 400), first, Watch::Diverged);
  // after an early minimum the residual may rise up to 1e3 x that minimum
  std::vector<double> osc = {1.0, 1e-2, 5.0, 9.0, 9.99, 3.0};
  expect("bounded rise above the minimum", run(osc, 1e-6, 5), 5, Watch::Go);
  std::vector<double> jump = {1.0, 1e-2, 5.0, 10.5};
  expect("rise past 1e3 x the minimum", run(jump, 1e-6, 10), 3, Watch::Diverged);
  expect("already converged", run({0.0, 1.0}, 0.0, 10), 0, Watch::Go);
  std::printf("%d wrong\n", bad);
  return bad ? 1 : 0;
}
