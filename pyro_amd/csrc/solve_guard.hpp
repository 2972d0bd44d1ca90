// solve_guard.hpp -- the residual watch of the iterative pressure solves
// (poisson.hip).  Plain C++ without HIP so that the host test
// tests/native/solve_guard_check.cpp compiles it with g++.
//
// Every relative residual a solve reads goes through ResidWatch::step:
//  - NaN or Inf stops the solve with an error (a NaN residual would pass
//    `rel > tol` as converged and never satisfy `rel <= tol`);
//  - a residual above kDivergeFactor times the best one so far stops it as
//    diverged (the reference V-cycle grows without bound on some hierarchies).
#ifndef PYRO_SOLVE_GUARD_HPP
#define PYRO_SOLVE_GUARD_HPP
#include <cmath>

namespace pyro {

constexpr double kDivergeFactor = 1e3;

enum class Watch { Go = 0, NonFinite = 1, Diverged = 2 };

struct ResidWatch {
  double best = 0.0;
  // k = number of iterations done (0 for the initial residual)
  Watch step(double rel, int k) {
    if (!std::isfinite(rel)) return Watch::NonFinite;
    if (k == 0 || rel < best) best = rel;
    if (rel > kDivergeFactor * best) return Watch::Diverged;
    return Watch::Go;
  }
};

}  // namespace pyro
#endif
