// Shared helpers of the functional programs: a failing check prints what
// failed and exits nonzero after op_exit.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>

#define FT_TOL 1e-9

inline void ft_check(bool ok, const char *what, int at) {
  if (!ok) {
    std::printf("FAILED: %s (at %d)\n", what, at);
    op_exit();
    std::exit(1);
  }
}

inline bool ft_close(double a, double b) { return std::fabs(a - b) < FT_TOL; }
