// The airfoil time-stepping program written against the OP2 API (sets, maps,
// dats, constants, five op_par_loops per step), for translation by
// op2amd.translator.  Usage: airfoil_tr <new_grid.dat> <iterations> [graph=G]
// [op_init tokens...].  Prints "iter rms maxerr errloc" per iteration with 17
// significant digits, the loop's wall time, op_timing_output's per-loop table
// (kernel times with the OP_HIP_KERNEL_TIMING token), then a checksum of q.  The step is the reference
// program's (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:221-297); the mesh
// file format is its reader's (:88-143).  With graph=G (built for the HIP
// backend) the first iteration runs eagerly and the rest as launches of a
// captured graph of up to G steps (a divisor of the remaining iterations),
// their rms/maxerr/errloc read from the graph history.  With dump=<file> the
// final q is also written with op_print_dat_to_binfile (airfoil.cpp:303).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "op_seq.h"

double gam, gm1, cfl, eps, mach, alpha, qinf[4];

#include "kernels.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::printf("usage: %s new_grid.dat iterations [op_init tokens]\n", argv[0]);
    return 2;
  }
  std::vector<char *> init_argv{argv[0]};
  int graph_steps = 0;
  const char *dump = nullptr;
  for (int i = 3; i < argc; ++i) {
    if (!std::strncmp(argv[i], "graph=", 6))
      graph_steps = std::atoi(argv[i] + 6);
    else if (!std::strncmp(argv[i], "dump=", 5))
      dump = argv[i] + 5;
    else
      init_argv.push_back(argv[i]);
  }
  op_init((int)init_argv.size(), init_argv.data(), 0);
  const int niter = std::atoi(argv[2]);

  FILE *fp = std::fopen(argv[1], "r");
  if (!fp) {
    std::printf("cannot open %s\n", argv[1]);
    return 2;
  }
  int nnode, ncell, nedge, nbedge;
  if (std::fscanf(fp, "%d %d %d %d", &nnode, &ncell, &nedge, &nbedge) != 4) return 2;
  std::vector<double> x(2 * nnode);
  std::vector<int> cell(4 * ncell), edge(2 * nedge), ecell(2 * nedge), bedge(2 * nbedge),
      becell(nbedge), bound(nbedge);
  for (int n = 0; n < nnode; ++n)
    if (std::fscanf(fp, "%lf %lf", &x[2 * n], &x[2 * n + 1]) != 2) return 2;
  for (int n = 0; n < ncell; ++n)
    if (std::fscanf(fp, "%d %d %d %d", &cell[4 * n], &cell[4 * n + 1], &cell[4 * n + 2],
                    &cell[4 * n + 3]) != 4)
      return 2;
  for (int n = 0; n < nedge; ++n)
    if (std::fscanf(fp, "%d %d %d %d", &edge[2 * n], &edge[2 * n + 1], &ecell[2 * n],
                    &ecell[2 * n + 1]) != 4)
      return 2;
  for (int n = 0; n < nbedge; ++n)
    if (std::fscanf(fp, "%d %d %d %d", &bedge[2 * n], &bedge[2 * n + 1], &becell[n],
                    &bound[n]) != 4)
      return 2;
  std::fclose(fp);

  gam = 1.4f;
  gm1 = gam - 1.0f;
  cfl = 0.9f;
  eps = 0.05f;
  mach = 0.4f;
  alpha = 3.0f * std::atan(1.0f) / 45.0f;  // 3 degrees, evaluated in float
  const double p = 1.0f, r = 1.0f;
  const double u = std::sqrt(gam * p / r) * mach;
  const double e = p / (r * gm1) + 0.5f * u * u;
  qinf[0] = r;
  qinf[1] = r * u * std::cos(alpha);
  qinf[2] = r * u * std::sin(alpha);
  qinf[3] = r * e;

  std::vector<double> q(4 * ncell), res(4 * ncell, 0.0), adt(ncell, 0.0);
  for (int n = 0; n < ncell; ++n)
    for (int k = 0; k < 4; ++k) q[4 * n + k] = qinf[k];

  op_set nodes = op_decl_set(nnode, "nodes");
  op_set edges = op_decl_set(nedge, "edges");
  op_set bedges = op_decl_set(nbedge, "bedges");
  op_set cells = op_decl_set(ncell, "cells");
  op_map pedge = op_decl_map(edges, nodes, 2, edge.data(), "pedge");
  op_map pecell = op_decl_map(edges, cells, 2, ecell.data(), "pecell");
  op_map pbedge = op_decl_map(bedges, nodes, 2, bedge.data(), "pbedge");
  op_map pbecell = op_decl_map(bedges, cells, 1, becell.data(), "pbecell");
  op_map pcell = op_decl_map(cells, nodes, 4, cell.data(), "pcell");
  op_dat p_bound = op_decl_dat(bedges, 1, "int", bound.data(), "p_bound");
  op_dat p_x = op_decl_dat(nodes, 2, "double", x.data(), "p_x");
  op_dat p_q = op_decl_dat(cells, 4, "double", q.data(), "p_q");
  op_dat p_qold = op_decl_dat(cells, 4, "double", q.data(), "p_qold");
  op_dat p_adt = op_decl_dat(cells, 1, "double", adt.data(), "p_adt");
  op_dat p_res = op_decl_dat(cells, 4, "double", res.data(), "p_res");

  op_decl_const(1, "double", &gam);
  op_decl_const(1, "double", &gm1);
  op_decl_const(1, "double", &cfl);
  op_decl_const(1, "double", &eps);
  op_decl_const(1, "double", &mach);
  op_decl_const(1, "double", &alpha);
  op_decl_const(4, "double", qinf);

  double cpu0, wall0, cpu1, wall1;
  op_timers(&cpu0, &wall0);
  double rms = 0.0, maxerr = 0.0;
  int errloc = 0;
  auto step = [&]() {
    op_par_loop(save_soln, "save_soln", cells,
                op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_READ),
                op_arg_dat(p_qold, -1, OP_ID, 4, "double", OP_WRITE));
    for (int k = 0; k < 2; k++) {
      op_par_loop(adt_calc, "adt_calc", cells,
                  op_arg_dat(p_x, 0, pcell, 2, "double", OP_READ),
                  op_arg_dat(p_x, 1, pcell, 2, "double", OP_READ),
                  op_arg_dat(p_x, 2, pcell, 2, "double", OP_READ),
                  op_arg_dat(p_x, 3, pcell, 2, "double", OP_READ),
                  op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_READ),
                  op_arg_dat(p_adt, -1, OP_ID, 1, "double", OP_WRITE));
      op_par_loop(res_calc, "res_calc", edges,
                  op_arg_dat(p_x, 0, pedge, 2, "double", OP_READ),
                  op_arg_dat(p_x, 1, pedge, 2, "double", OP_READ),
                  op_arg_dat(p_q, 0, pecell, 4, "double", OP_READ),
                  op_arg_dat(p_q, 1, pecell, 4, "double", OP_READ),
                  op_arg_dat(p_adt, 0, pecell, 1, "double", OP_READ),
                  op_arg_dat(p_adt, 1, pecell, 1, "double", OP_READ),
                  op_arg_dat(p_res, 0, pecell, 4, "double", OP_INC),
                  op_arg_dat(p_res, 1, pecell, 4, "double", OP_INC));
      op_par_loop(bres_calc, "bres_calc", bedges,
                  op_arg_dat(p_x, 0, pbedge, 2, "double", OP_READ),
                  op_arg_dat(p_x, 1, pbedge, 2, "double", OP_READ),
                  op_arg_dat(p_q, 0, pbecell, 4, "double", OP_READ),
                  op_arg_dat(p_adt, 0, pbecell, 1, "double", OP_READ),
                  op_arg_dat(p_res, 0, pbecell, 4, "double", OP_INC),
                  op_arg_dat(p_bound, -1, OP_ID, 1, "int", OP_READ));
      rms = 0.0;
      maxerr = 0.0;
      errloc = 0;
      op_par_loop(update, "update", cells,
                  op_arg_dat(p_qold, -1, OP_ID, 4, "double", OP_READ),
                  op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_WRITE),
                  op_arg_dat(p_res, -1, OP_ID, 4, "double", OP_RW),
                  op_arg_dat(p_adt, -1, OP_ID, 1, "double", OP_READ),
                  op_arg_gbl(&rms, 1, "double", OP_INC),
                  op_arg_gbl(&maxerr, 1, "double", OP_MAX),
                  op_arg_idx(-1, OP_ID),
                  op_arg_info(&errloc, 1, "int", 5));
    }
  };
  if (graph_steps <= 0 || niter < 2) {
    for (int iter = 1; iter <= niter; iter++) {
      step();
      std::printf("%d %.17e %.17e %d\n", iter, std::sqrt(rms / ncell), maxerr, errloc);
    }
  } else {
#ifdef OP2_HIP_H
    // the first step eagerly (plans are built at each loop's first call),
    // then graph launches of G steps; each launch appends its 2 x G updates'
    // rms, maxerr, errloc (one 8-byte word each) to the device history
    step();
    std::printf("%d %.17e %.17e %d\n", 1, std::sqrt(rms / ncell), maxerr, errloc);
    const int rest = niter - 1;
    int G = graph_steps < rest ? graph_steps : rest;
    while (rest % G) --G;
    if (op2h_graph_begin() != 0) return 3;
    for (int k = 0; k < G; ++k) step();
    if (op2h_graph_end() != 0) return 3;
    const int launches = rest / G;
    op2h_graph_launch(launches);
    op2h_device_sync();
    int nrec = 0;
    const int words = op2h_graph_history(0, nullptr, &nrec);  // the record length
    if (words != 6 * G) {
      std::printf("graph history: records of %d words\n", words);
      return 4;
    }
    std::vector<double> hist((size_t)launches * words);
    op2h_graph_history(launches, hist.data(), &nrec);
    if (nrec != launches) {
      std::printf("graph history: %d records of %d words\n", nrec, words);
      return 4;
    }
    for (int i = 0; i < rest; ++i) {
      const double *h = &hist[(size_t)i * 6 + 3];  // the step's second update
      // delivered as (host value at capture, 0) op (device result)
      const double mx = h[1] > 0.0 ? h[1] : 0.0;
      int loc = 0;
      if (h[1] > 0.0) std::memcpy(&loc, &h[2], sizeof(int));
      std::printf("%d %.17e %.17e %d\n", i + 2, std::sqrt(h[0] / ncell), mx, loc);
    }
    op2h_graph_sync();
    op2h_graph_destroy();
#else
    std::printf("graph=G needs the HIP backend\n");
    return 2;
#endif
  }
  op_timers(&cpu1, &wall1);
  std::printf("time %.6f s for %d iterations\n", wall1 - wall0, niter);
  op_timing_output();
  std::vector<double> qout(4 * ncell);
  op_fetch_data(p_q, qout.data());
  double sum = 0.0, sumsq = 0.0;
  for (double v : qout) {
    sum += v;
    sumsq += v * v;
  }
  std::printf("q %.17e %.17e\n", sum, sumsq);
  if (dump) op_print_dat_to_binfile(p_q, dump);
  op_exit();
  return 0;
}
