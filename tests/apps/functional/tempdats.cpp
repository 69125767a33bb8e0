// Temporary dats (op_decl_dat_temp / op_free_dat_temp, op_lib_core.cpp:643-680
// of the reference): a temporary dat starts zeroed, increments through a map
// land in it like in any dat, op_free_dat_temp returns 1, and a dat declared
// again under the same name after the free starts from zero again.  Repeated
// 40 times, as an application that declares its work arrays every iteration
// (apps/c/airfoil/airfoil_tempdats/dp/airfoil.cpp:225-307) does.  Expected
// values are computed on the host from the inputs.
#include <vector>

#include "op_seq.h"
#include "check.h"

static void edge_to_nodes(const double *w, double *a, double *b) {
  a[0] += w[0];
  a[1] += 2.0 * w[0];
  b[0] += w[0];
  b[1] -= w[0];
}
static void scale_into(const double *t, const double *s, double *out) {
  out[0] = s[0] * (t[0] + t[1]);
}

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  // a ring of 96 nodes with chords: every node is reached by four edges
  const int nn = 96, ne = 2 * nn;
  std::vector<int> e2n(2 * ne);
  for (int e = 0; e < nn; ++e) {
    e2n[2 * e] = e;
    e2n[2 * e + 1] = (e + 1) % nn;
    e2n[2 * (nn + e)] = (e + 29) % nn;
    e2n[2 * (nn + e) + 1] = e;
  }
  std::vector<double> w(ne), s(nn);
  for (int e = 0; e < ne; ++e) w[e] = 0.25 + 0.125 * (e % 9);
  for (int i = 0; i < nn; ++i) s[i] = 1.0 + 0.5 * (i % 5);
  op_set nodes = op_decl_set(nn, "nodes");
  op_set edges = op_decl_set(ne, "edges");
  op_map m = op_decl_map(edges, nodes, 2, e2n.data(), "edge_nodes");
  op_dat dw = op_decl_dat(edges, 1, "double", w.data(), "w");
  op_dat ds = op_decl_dat(nodes, 1, "double", s.data(), "s");
  std::vector<double> zero(nn, 0.0);
  op_dat dout = op_decl_dat(nodes, 1, "double", zero.data(), "out");

  // host expectation of one pass
  std::vector<double> t(2 * nn, 0.0), want(nn);
  for (int e = 0; e < ne; ++e) {
    const int a = e2n[2 * e], b = e2n[2 * e + 1];
    t[2 * a] += w[e];
    t[2 * a + 1] += 2.0 * w[e];
    t[2 * b] += w[e];
    t[2 * b + 1] -= w[e];
  }
  for (int i = 0; i < nn; ++i) want[i] = s[i] * (t[2 * i] + t[2 * i + 1]);

  std::vector<double> got(2 * nn), out(nn);
  for (int it = 0; it < 40; ++it) {
    double *none = nullptr;
    op_dat tmp = op_decl_dat_temp(nodes, 2, "double", none, "t");
    if (it == 0) {
      op_fetch_data(tmp, got.data());
      for (int i = 0; i < 2 * nn; ++i) ft_check(got[i] == 0.0, "temporary dat starts zeroed", i);
    }
    op_par_loop(edge_to_nodes, "edge_to_nodes", edges,
                op_arg_dat(dw, -1, OP_ID, 1, "double", OP_READ),
                op_arg_dat(tmp, 0, m, 2, "double", OP_INC),
                op_arg_dat(tmp, 1, m, 2, "double", OP_INC));
    op_par_loop(scale_into, "scale_into", nodes, op_arg_dat(tmp, -1, OP_ID, 2, "double", OP_READ),
                op_arg_dat(ds, -1, OP_ID, 1, "double", OP_READ),
                op_arg_dat(dout, -1, OP_ID, 1, "double", OP_WRITE));
    op_fetch_data(tmp, got.data());
    for (int i = 0; i < 2 * nn; ++i)
      ft_check(ft_close(got[i], t[i]), "increments into a temporary dat (one pass)", i);
    op_fetch_data(dout, out.data());
    for (int i = 0; i < nn; ++i) ft_check(ft_close(out[i], want[i]), "read of a temporary dat", i);
    ft_check(op_free_dat_temp(tmp) == 1, "op_free_dat_temp returns 1", it);
  }
  std::printf("tempdats_cycle passed\n");

  op_exit();
  std::printf("tempdats: all passed\n");
  return 0;
}
