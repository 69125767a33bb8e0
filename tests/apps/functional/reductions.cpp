// Increments of dats: through a map (each edge adds to both of its nodes,
// dims 1 and 3, several edges per node) and directly (dims 1 and 4).
// Coverage follows tests/functional/dat_reductions/reduc_tests.cpp:36-200 of
// the reference; expected values are computed on the host from the inputs.
#include <vector>

#include "op_seq.h"
#include "check.h"

static void edge_to_nodes1(float *a, float *b, const float *w) {
  a[0] += w[0];
  b[0] += w[0];
}
static void edge_to_nodes3(float *a, float *b, const float *w) {
  for (int k = 0; k < 3; ++k) {
    a[k] += w[k];
    b[k] += w[k];
  }
}
static void self_inc1(const float *w, float *acc) { acc[0] += w[0] + 3.25f; }
static void self_inc4(const float *w, float *acc) {
  for (int k = 0; k < 4; ++k) acc[k] += w[k] + 1.325f * k;
}

int main(int argc, char **argv) {
  op_init(argc, argv, 1);
  // a ring of 144 nodes; edge e joins e and (e + 1) mod 144, and a chord
  // joins e and (e + 37) mod 144, so every node is reached by four edges
  const int nn = 144, ne = 2 * nn;
  std::vector<int> e2n(2 * ne);
  for (int e = 0; e < nn; ++e) {
    e2n[2 * e] = e;
    e2n[2 * e + 1] = (e + 1) % nn;
    e2n[2 * (nn + e)] = (e + 37) % nn;
    e2n[2 * (nn + e) + 1] = e;
  }
  std::vector<float> ew1(ne), ew4(4 * ne), n1(nn), n3(3 * nn);
  for (int e = 0; e < ne; ++e) {
    ew1[e] = 0.5f + 0.25f * (e % 7);
    for (int k = 0; k < 4; ++k) ew4[4 * e + k] = 0.125f * ((e + k) % 11);
  }
  for (int i = 0; i < nn; ++i) {
    n1[i] = 1.0f + i;
    for (int k = 0; k < 3; ++k) n3[3 * i + k] = 2.0f * i + k;
  }
  op_set nodes = op_decl_set(nn, "nodes");
  op_set edges = op_decl_set(ne, "edges");
  op_map m = op_decl_map(edges, nodes, 2, e2n.data(), "edge_nodes");
  op_dat w1 = op_decl_dat(edges, 1, "float", ew1.data(), "w1");
  op_dat w4 = op_decl_dat(edges, 4, "float", ew4.data(), "w4");
  op_dat a1 = op_decl_dat(edges, 1, "float", ew1.data(), "a1");
  op_dat a4 = op_decl_dat(edges, 4, "float", ew4.data(), "a4");
  op_dat v1 = op_decl_dat(nodes, 1, "float", n1.data(), "v1");
  op_dat v3 = op_decl_dat(nodes, 3, "float", n3.data(), "v3");

  op_par_loop(edge_to_nodes1, "edge_to_nodes1", edges,
              op_arg_dat(v1, 0, m, 1, "float", OP_INC), op_arg_dat(v1, 1, m, 1, "float", OP_INC),
              op_arg_dat(w1, -1, OP_ID, 1, "float", OP_READ));
  std::vector<float> got(4 * ne);
  std::vector<double> want(n1.begin(), n1.end());
  for (int e = 0; e < ne; ++e) {
    want[e2n[2 * e]] += ew1[e];
    want[e2n[2 * e + 1]] += ew1[e];
  }
  op_fetch_data(v1, got.data());
  for (int i = 0; i < nn; ++i) ft_check(std::fabs(got[i] - want[i]) < 1e-4, "indirect inc dim 1", i);
  std::printf("indirect_inc1 passed\n");

  op_par_loop(edge_to_nodes3, "edge_to_nodes3", edges,
              op_arg_dat(v3, 0, m, 3, "float", OP_INC), op_arg_dat(v3, 1, m, 3, "float", OP_INC),
              op_arg_dat(w4, -1, OP_ID, 4, "float", OP_READ));
  std::vector<double> want3(n3.begin(), n3.end());
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) {
      want3[3 * e2n[2 * e] + k] += ew4[4 * e + k];
      want3[3 * e2n[2 * e + 1] + k] += ew4[4 * e + k];
    }
  op_fetch_data(v3, got.data());
  for (int i = 0; i < 3 * nn; ++i)
    ft_check(std::fabs(got[i] - want3[i]) < 1e-4, "indirect inc dim 3", i);
  std::printf("indirect_inc3 passed\n");

  // many increments per target: 600 spokes join one of three hubs (nodes 0,
  // 48, 96) to node s mod 144, so a hub takes over 200 increments and each
  // block of the owner partition walks its spokes in several chunks; the
  // weights are multiples of 1/4, so every sum is exact in float
  const int ns = 600;
  std::vector<int> s2n(2 * ns);
  std::vector<float> sw(ns);
  for (int e = 0; e < ns; ++e) {
    s2n[2 * e] = 48 * (e % 3);
    s2n[2 * e + 1] = e % nn;
    sw[e] = 0.25f * (e % 5);
  }
  op_set spokes = op_decl_set(ns, "spokes");
  op_map ms = op_decl_map(spokes, nodes, 2, s2n.data(), "spoke_nodes");
  op_dat swd = op_decl_dat(spokes, 1, "float", sw.data(), "sw");
  op_fetch_data(v1, got.data());
  std::vector<double> wants(got.begin(), got.begin() + nn);
  for (int e = 0; e < ns; ++e) {
    wants[s2n[2 * e]] += sw[e];
    wants[s2n[2 * e + 1]] += sw[e];
  }
  op_par_loop(edge_to_nodes1, "spokes_to_nodes", spokes,
              op_arg_dat(v1, 0, ms, 1, "float", OP_INC), op_arg_dat(v1, 1, ms, 1, "float", OP_INC),
              op_arg_dat(swd, -1, OP_ID, 1, "float", OP_READ));
  op_fetch_data(v1, got.data());
  for (int i = 0; i < nn; ++i)
    ft_check(got[i] == (float)wants[i], "many increments per target", i);
  std::printf("hub_inc passed\n");

  // targets far apart in numbering: on 140000 nodes, 500 links join
  // 137e mod 140000 to a node 70001 further on, 500 more join neighbours, so
  // some blocks of the owner partition span more than 65535 node numbers and
  // others do not
  const int nf = 140000, nl = 1000;
  std::vector<int> l2n(2 * nl);
  std::vector<float> lw(nl), fv(nf, 0.5f);
  for (int e = 0; e < nl; ++e) {
    const int a = e < 500 ? (137 * e) % nf : 90000 + e;
    l2n[2 * e] = a;
    l2n[2 * e + 1] = e < 500 ? (a + 70001) % nf : a + 1;
    lw[e] = 0.25f * (1 + e % 3);
  }
  op_set far_nodes = op_decl_set(nf, "far_nodes");
  op_set links = op_decl_set(nl, "links");
  op_map ml = op_decl_map(links, far_nodes, 2, l2n.data(), "link_nodes");
  op_dat lwd = op_decl_dat(links, 1, "float", lw.data(), "lw");
  op_dat fvd = op_decl_dat(far_nodes, 1, "float", fv.data(), "fv");
  std::vector<double> wantf(fv.begin(), fv.end());
  for (int e = 0; e < nl; ++e) {
    wantf[l2n[2 * e]] += lw[e];
    wantf[l2n[2 * e + 1]] += lw[e];
  }
  op_par_loop(edge_to_nodes1, "links_to_nodes", links,
              op_arg_dat(fvd, 0, ml, 1, "float", OP_INC), op_arg_dat(fvd, 1, ml, 1, "float", OP_INC),
              op_arg_dat(lwd, -1, OP_ID, 1, "float", OP_READ));
  std::vector<float> gotf(nf);
  op_fetch_data(fvd, gotf.data());
  for (int i = 0; i < nf; ++i)
    ft_check(gotf[i] == (float)wantf[i], "targets far apart in numbering", i);
  std::printf("far_inc passed\n");

  op_par_loop(self_inc1, "self_inc1", edges, op_arg_dat(w1, -1, OP_ID, 1, "float", OP_READ),
              op_arg_dat(a1, -1, OP_ID, 1, "float", OP_INC));
  op_fetch_data(a1, got.data());
  for (int e = 0; e < ne; ++e)
    ft_check(std::fabs(got[e] - (2.0f * ew1[e] + 3.25f)) < 1e-5, "direct inc dim 1", e);
  std::printf("direct_inc1 passed\n");

  op_par_loop(self_inc4, "self_inc4", edges, op_arg_dat(w4, -1, OP_ID, 4, "float", OP_READ),
              op_arg_dat(a4, -1, OP_ID, 4, "float", OP_INC));
  op_fetch_data(a4, got.data());
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 4; ++k)
      ft_check(std::fabs(got[4 * e + k] - (2.0f * ew4[4 * e + k] + 1.325f * k)) < 1e-5,
               "direct inc dim 4", 4 * e + k);
  std::printf("direct_inc4 passed\n");

  op_exit();
  std::printf("reductions: all passed\n");
  return 0;
}
