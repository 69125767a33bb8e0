// User kernels that call other functions of the program: a helper calling a
// helper, a helper from a quoted include, and a helper shared by two kernels.
// The translator copies each called function to the device once, callees
// before callers.
#include <vector>

#include "op_seq.h"
#include "check.h"
#include "helpers_lib.h"

static inline double square(double v) { return v * v; }
static inline double poly(double v) { return square(v) + 2.0 * cube(v) + 1.0; }

static void k_poly(const double *x, double *y) { y[0] = poly(x[0]); }
static void k_squares(const double *x, double *y) { y[0] = square(x[0]) + square(2.0 * x[0]); }

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  const int n = 50;
  std::vector<double> x(n), z(n, 0.0), got(n);
  for (int i = 0; i < n; ++i) x[i] = 0.1 * i - 2.0;
  op_set set = op_decl_set(n, "elements");
  op_dat dx = op_decl_dat(set, 1, "double", x.data(), "x");
  op_dat dy = op_decl_dat(set, 1, "double", z.data(), "y");

  op_par_loop(k_poly, "k_poly", set, op_arg_dat(dx, -1, OP_ID, 1, "double", OP_READ),
              op_arg_dat(dy, -1, OP_ID, 1, "double", OP_WRITE));
  op_fetch_data(dy, got.data());
  for (int i = 0; i < n; ++i)
    ft_check(ft_close(got[i], x[i] * x[i] + 2.0 * x[i] * x[i] * x[i] + 1.0), "helper chain", i);
  std::printf("helpers_chain passed\n");

  op_par_loop(k_squares, "k_squares", set, op_arg_dat(dx, -1, OP_ID, 1, "double", OP_READ),
              op_arg_dat(dy, -1, OP_ID, 1, "double", OP_WRITE));
  op_fetch_data(dy, got.data());
  for (int i = 0; i < n; ++i) ft_check(ft_close(got[i], 5.0 * x[i] * x[i]), "shared helper", i);
  std::printf("helpers_shared passed\n");

  op_exit();
  std::printf("helpers: all passed\n");
  return 0;
}
