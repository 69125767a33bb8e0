// op_arg_idx: the element's own index, and the target index through each
// column of a map, alone and mixed with dat arguments.  Coverage follows
// tests/functional/idx/idx_tests.cpp:17-168 of the reference.
#include <vector>

#include "op_seq.h"
#include "check.h"

static void own_index(double *out, const int *i) { out[0] = (double)i[0]; }
static void map_index(double *out, const int *a, const int *b, const int *c) {
  out[0] = (double)a[0];
  out[1] = (double)b[0];
  out[2] = (double)c[0];
}
static void both(double *out, const int *self, const int *a, const int *b, const int *c) {
  out[0] = (double)self[0];
  out[1] = (double)a[0];
  out[2] = (double)b[0];
  out[3] = (double)c[0];
}

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  const int ne = 48, nn = 64, md = 3;
  std::vector<int> p(ne * md), q(ne * md);
  for (int e = 0; e < ne; ++e)
    for (int d = 0; d < md; ++d) {
      p[e * md + d] = (e * 7 + d * 11) % nn;
      q[e * md + d] = (e * 5 + d * 3 + 1) % nn;
    }
  op_set edges = op_decl_set(ne, "edges");
  op_set nodes = op_decl_set(nn, "nodes");
  op_map mp = op_decl_map(edges, nodes, md, p.data(), "p");
  op_map mq = op_decl_map(edges, nodes, md, q.data(), "q");
  std::vector<double> z(4 * ne, 0.0), got(4 * ne);
  op_dat o1 = op_decl_dat(edges, 1, "double", z.data(), "o1");
  op_dat o3 = op_decl_dat(edges, 3, "double", z.data(), "o3");
  op_dat o4 = op_decl_dat(edges, 4, "double", z.data(), "o4");

  op_par_loop(own_index, "own_index", edges, op_arg_dat(o1, -1, OP_ID, 1, "double", OP_WRITE),
              op_arg_idx(-1, OP_ID));
  op_fetch_data(o1, got.data());
  for (int e = 0; e < ne; ++e) ft_check(got[e] == e, "direct idx", e);
  std::printf("direct_idx passed\n");

  op_par_loop(map_index, "map_index", edges, op_arg_dat(o3, -1, OP_ID, 3, "double", OP_WRITE),
              op_arg_idx(0, mp), op_arg_idx(1, mp), op_arg_idx(2, mp));
  op_fetch_data(o3, got.data());
  for (int e = 0; e < ne; ++e)
    for (int d = 0; d < md; ++d) ft_check(got[e * 3 + d] == p[e * md + d], "map idx", e * 3 + d);
  std::printf("indirect_idx passed\n");

  op_par_loop(both, "both", edges, op_arg_dat(o4, -1, OP_ID, 4, "double", OP_WRITE),
              op_arg_idx(-1, OP_ID), op_arg_idx(0, mq), op_arg_idx(2, mq), op_arg_idx(1, mq));
  op_fetch_data(o4, got.data());
  for (int e = 0; e < ne; ++e) {
    ft_check(got[4 * e] == e, "mixed idx: own", e);
    ft_check(got[4 * e + 1] == q[e * md + 0], "mixed idx: column 0", e);
    ft_check(got[4 * e + 2] == q[e * md + 2], "mixed idx: column 2", e);
    ft_check(got[4 * e + 3] == q[e * md + 1], "mixed idx: column 1", e);
  }
  std::printf("mixed_idx passed\n");

  op_exit();
  std::printf("idx: all passed\n");
  return 0;
}
