// Global arguments: OP_READ values reach every element, OP_INC/OP_MIN/OP_MAX
// reduce over the set, op_arg_info reports the first element reaching the
// extremum.  Coverage follows tests/functional/gbl/gbl_tests.cpp:79-220 of the
// reference (dims 1 and 5 of each access) on a set whose size is not a
// multiple of a wave or a workgroup.
#include <numeric>
#include <vector>

#include "op_seq.h"
#include "check.h"

static void copy_g1(double *out, const double *g) { out[0] = g[0]; }
static void copy_g5(double *out, const double *g) {
  for (int k = 0; k < 5; ++k) out[k] = g[k];
}
static void sum1(const double *v, double *s) { s[0] += v[0]; }
static void sum5(const double *v, double *s) {
  for (int k = 0; k < 5; ++k) s[k] += v[k];
}
static void low1(const double *v, double *m) { m[0] = MIN(m[0], v[0]); }
static void low5(const double *v, double *m) {
  for (int k = 0; k < 5; ++k) m[k] = MIN(m[k], v[k]);
}
static void high1(const double *v, double *m) { m[0] = MAX(m[0], v[0]); }
static void high5(const double *v, double *m) {
  for (int k = 0; k < 5; ++k) m[k] = MAX(m[k], v[k]);
}
static void count_int(const double *v, int *c) { c[0] += v[0] >= 0.0 ? 1 : 0; }
static void argmax(const double *v, const int *i, double *m, int *where) {
  if (v[0] > m[0]) {
    m[0] = v[0];
    where[0] = i[0];
  }
}

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  const int n = 161;
  std::vector<double> a(n), b(5 * n), zeros(5 * n, 0.0), peaks(n);
  std::iota(a.begin(), a.end(), 0.0);
  std::iota(b.begin(), b.end(), 0.0);
  int first = 0;  // the first element holding the largest peak
  for (int i = 0; i < n; ++i) {
    peaks[i] = (double)((i * 37) % 50);
    if (peaks[i] > peaks[first]) first = i;
  }
  op_set set = op_decl_set(n, "elements");
  op_dat d1 = op_decl_dat(set, 1, "double", a.data(), "d1");
  op_dat d5 = op_decl_dat(set, 5, "double", b.data(), "d5");
  op_dat o1 = op_decl_dat(set, 1, "double", zeros.data(), "o1");
  op_dat o5 = op_decl_dat(set, 5, "double", zeros.data(), "o5");
  op_dat pk = op_decl_dat(set, 1, "double", peaks.data(), "peaks");
  std::vector<double> got(5 * n);

  double g1 = 20.0;
  op_par_loop(copy_g1, "copy_g1", set, op_arg_dat(o1, -1, OP_ID, 1, "double", OP_WRITE),
              op_arg_gbl(&g1, 1, "double", OP_READ));
  op_fetch_data(o1, got.data());
  for (int i = 0; i < n; ++i) ft_check(ft_close(got[i], 20.0), "gbl read dim 1", i);
  std::printf("read1 passed\n");

  double g5[5] = {30.0, 40.0, 50.0, 60.0, 70.0};
  op_par_loop(copy_g5, "copy_g5", set, op_arg_dat(o5, -1, OP_ID, 5, "double", OP_WRITE),
              op_arg_gbl(g5, 5, "double", OP_READ));
  op_fetch_data(o5, got.data());
  for (int i = 0; i < 5 * n; ++i) ft_check(ft_close(got[i], g5[i % 5]), "gbl read dim 5", i);
  std::printf("read5 passed\n");

  double s1 = 1.5;  // the loop adds to what the caller holds
  op_par_loop(sum1, "sum1", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_READ),
              op_arg_gbl(&s1, 1, "double", OP_INC));
  ft_check(ft_close(s1, 1.5 + n * (n - 1) / 2.0), "gbl inc dim 1", 0);
  std::printf("inc1 passed\n");

  double s5[5] = {0, 0, 0, 0, 0};
  op_par_loop(sum5, "sum5", set, op_arg_dat(d5, -1, OP_ID, 5, "double", OP_READ),
              op_arg_gbl(s5, 5, "double", OP_INC));
  for (int k = 0; k < 5; ++k) {
    double want = 0.0;  // sum over i of 5 i + k
    for (int i = 0; i < n; ++i) want += 5.0 * i + k;
    ft_check(ft_close(s5[k], want), "gbl inc dim 5", k);
  }
  std::printf("inc5 passed\n");

  double m1 = 1e300;
  op_par_loop(low1, "low1", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_READ),
              op_arg_gbl(&m1, 1, "double", OP_MIN));
  ft_check(ft_close(m1, 0.0), "gbl min dim 1", 0);
  double m1b = -3.0;  // already below every element: unchanged
  op_par_loop(low1, "low1", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_READ),
              op_arg_gbl(&m1b, 1, "double", OP_MIN));
  ft_check(ft_close(m1b, -3.0), "gbl min keeps a smaller initial value", 0);
  std::printf("min1 passed\n");

  double m5[5] = {1e300, 1e300, 1e300, 2.5, -1.0};
  op_par_loop(low5, "low5", set, op_arg_dat(d5, -1, OP_ID, 5, "double", OP_READ),
              op_arg_gbl(m5, 5, "double", OP_MIN));
  const double min5_want[5] = {0.0, 1.0, 2.0, 2.5, -1.0};
  for (int k = 0; k < 5; ++k) ft_check(ft_close(m5[k], min5_want[k]), "gbl min dim 5", k);
  std::printf("min5 passed\n");

  double x1 = -1e300;
  op_par_loop(high1, "high1", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_READ),
              op_arg_gbl(&x1, 1, "double", OP_MAX));
  ft_check(ft_close(x1, n - 1.0), "gbl max dim 1", 0);
  std::printf("max1 passed\n");

  double x5[5] = {-1e300, -1e300, -1e300, -1e300, 1e6};
  op_par_loop(high5, "high5", set, op_arg_dat(d5, -1, OP_ID, 5, "double", OP_READ),
              op_arg_gbl(x5, 5, "double", OP_MAX));
  for (int k = 0; k < 4; ++k) ft_check(ft_close(x5[k], 5.0 * (n - 1) + k), "gbl max dim 5", k);
  ft_check(ft_close(x5[4], 1e6), "gbl max keeps a larger initial value", 4);
  std::printf("max5 passed\n");

  int cnt = 7;
  op_par_loop(count_int, "count_int", set, op_arg_dat(d1, -1, OP_ID, 1, "double", OP_READ),
              op_arg_gbl(&cnt, 1, "int", OP_INC));
  ft_check(cnt == 7 + n, "gbl inc int", 0);
  std::printf("inc_int passed\n");

  double best = -1.0;
  int where = -5;
  op_par_loop(argmax, "argmax", set, op_arg_dat(pk, -1, OP_ID, 1, "double", OP_READ),
              op_arg_idx(-1, OP_ID), op_arg_gbl(&best, 1, "double", OP_MAX),
              op_arg_info(&where, 1, "int", 2));
  ft_check(ft_close(best, peaks[first]), "gbl max with info: value", 0);
  ft_check(where == first, "gbl max with info: first element reaching the maximum", where);
  std::printf("maxinfo passed\n");

  op_exit();
  std::printf("gbl: all passed\n");
  return 0;
}
