// Constants: a scalar and an array declared with op_decl_const are visible in
// kernels; a second declaration updates them.  Coverage follows
// tests/functional/const/const_tests.cpp:17-79 of the reference.
#include <vector>

#include "op_seq.h"
#include "check.h"

double scale;
double offsets[4];

static void use_scalar(double *out) { out[0] = scale; }
static void use_array(double *out) {
  for (int k = 0; k < 4; ++k) out[k] = offsets[k] * scale;
}

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  const int n = 70;
  std::vector<double> z(4 * n, 0.0), got(4 * n);
  op_set set = op_decl_set(n, "elements");
  op_dat d1 = op_decl_dat(set, 1, "double", z.data(), "d1");
  op_dat d4 = op_decl_dat(set, 4, "double", z.data(), "d4");
  scale = 2.5;
  offsets[0] = 1.0;
  offsets[1] = -2.0;
  offsets[2] = 0.25;
  offsets[3] = 8.0;
  op_decl_const(1, "double", &scale);
  op_decl_const(4, "double", offsets);

  op_par_loop(use_scalar, "use_scalar", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_WRITE));
  op_fetch_data(d1, got.data());
  for (int i = 0; i < n; ++i) ft_check(ft_close(got[i], 2.5), "scalar constant", i);
  std::printf("const_scalar passed\n");

  op_par_loop(use_array, "use_array", set, op_arg_dat(d4, -1, OP_ID, 4, "double", OP_WRITE));
  op_fetch_data(d4, got.data());
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 4; ++k) ft_check(ft_close(got[4 * i + k], offsets[k] * 2.5), "array constant", 4 * i + k);
  std::printf("const_array passed\n");

  scale = -1.0;
  op_decl_const(1, "double", &scale);
  op_par_loop(use_scalar, "use_scalar", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_WRITE));
  op_fetch_data(d1, got.data());
  for (int i = 0; i < n; ++i) ft_check(ft_close(got[i], -1.0), "redeclared constant", i);
  std::printf("const_update passed\n");

  op_exit();
  std::printf("consts: all passed\n");
  return 0;
}
