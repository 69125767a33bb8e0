// Dats in structure-of-arrays layout on the device ("double:soa", and every
// multi-component dat under OP_AUTO_SOA): direct writes and reads of dim 5,
// two strided dats written by one kernel, and an indirect read through a map
// of a strided dat.  Coverage follows tests/functional/strides/
// stride_tests.cpp:21-120 of the reference.
#include <vector>

#include "op_seq.h"
#include "check.h"

static void fill5(double *out, const double *base) {
  for (int k = 0; k < 5; ++k) out[k] = base[0] * 10.0 + k;
}
static void fill_two(double *a, double *b, const double *base) {
  for (int k = 0; k < 5; ++k) {
    a[k] = base[0] + 0.5 * k;
    b[k] = -base[0] - 0.25 * k;
  }
}
static void gather_next(double *out, const double *nxt) {
  for (int k = 0; k < 5; ++k) out[k] = nxt[k] + 1.0;
}

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  const int n = 97;
  std::vector<double> base(n), z(5 * n, 0.0), got(5 * n);
  for (int i = 0; i < n; ++i) base[i] = (double)i;
  std::vector<int> next(n);
  for (int i = 0; i < n; ++i) next[i] = (i + 3) % n;
  op_set set = op_decl_set(n, "elements");
  op_map mnext = op_decl_map(set, set, 1, next.data(), "next");
  op_dat b = op_decl_dat(set, 1, "double", base.data(), "base");
  op_dat s0 = op_decl_dat(set, 5, "double:soa", z.data(), "s0");
  op_dat s1 = op_decl_dat(set, 5, "double", z.data(), "s1");
  op_dat s2 = op_decl_dat(set, 5, "double:soa", z.data(), "s2");

  op_par_loop(fill5, "fill5", set, op_arg_dat(s0, -1, OP_ID, 5, "double", OP_WRITE),
              op_arg_dat(b, -1, OP_ID, 1, "double", OP_READ));
  op_fetch_data(s0, got.data());
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 5; ++k) ft_check(ft_close(got[5 * i + k], 10.0 * i + k), "soa write", 5 * i + k);
  std::printf("soa_write passed\n");

  op_par_loop(fill_two, "fill_two", set, op_arg_dat(s1, -1, OP_ID, 5, "double", OP_WRITE),
              op_arg_dat(s2, -1, OP_ID, 5, "double", OP_WRITE),
              op_arg_dat(b, -1, OP_ID, 1, "double", OP_READ));
  op_fetch_data(s1, got.data());
  for (int i = 0; i < 5 * n; ++i) ft_check(ft_close(got[i], (i / 5) + 0.5 * (i % 5)), "two dats: first", i);
  op_fetch_data(s2, got.data());
  for (int i = 0; i < 5 * n; ++i) ft_check(ft_close(got[i], -(i / 5) - 0.25 * (i % 5)), "two dats: second", i);
  std::printf("soa_two_dats passed\n");

  op_par_loop(gather_next, "gather_next", set, op_arg_dat(s1, -1, OP_ID, 5, "double", OP_WRITE),
              op_arg_dat(s0, 0, mnext, 5, "double", OP_READ));
  op_fetch_data(s1, got.data());
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 5; ++k)
      ft_check(ft_close(got[5 * i + k], 10.0 * next[i] + k + 1.0), "soa indirect read", 5 * i + k);
  std::printf("soa_indirect_read passed\n");

  op_exit();
  std::printf("strides: all passed\n");
  return 0;
}
