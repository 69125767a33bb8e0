// Test-infrastructure harness (oracle only; never shipped in the product).
//
// Drives the REFERENCE's own sequential executor (op2/include/op_seq.h:918-1073)
// and the reference airfoil kernels (apps/c/airfoil/airfoil_plain/dp/*.h) on a
// mesh in the new_grid.dat format, and records what the reference app only
// prints every 100 iterations (airfoil.cpp:280-283): per-iteration rms, maxerr
// and errloc with full precision, the final q, and the state after the first
// loop calls of iteration 1.  Built by oracle/Makefile against the sources
// where they lie under the reference tree; outputs go to oracle/_ref/.
//
// Usage: ref_harness GRID NITER ALPHA_MODE OUTPREFIX
//   ALPHA_MODE 0: qinf[2] = 0 exactly as airfoil.cpp:161-164 (alpha unused)
//   ALPHA_MODE 1: free stream rotated by alpha = 3.0f*atan(1.0f)/45.0f:
//                 qinf[1] = r*u*cos(alpha), qinf[2] = r*u*sin(alpha)
// Outputs:
//   OUTPREFIX_hist.txt  "iter rms_raw_k0 rms_raw_k1 rms maxerr errloc" per iteration
//   OUTPREFIX_q.bin     final q, ncell*4 doubles (AoS)
//   OUTPREFIX_first.bin iteration 1: adt after adt_calc, res after res_calc,
//                       res after bres_calc, q after update k=0 (all doubles)
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

// globals referenced by the reference kernel headers
double gam, gm1, cfl, eps, mach, alpha, qinf[4];

#include "op_seq.h"

#include "adt_calc.h"
#include "bres_calc.h"
#include "res_calc.h"
#include "save_soln.h"
#include "update.h"

static void die(const char *msg) {
  fprintf(stderr, "ref_harness: %s\n", msg);
  exit(1);
}

static void put(FILE *fp, const double *p, size_t n) {
  if (fwrite(p, sizeof(double), n, fp) != n) die("write failed");
}

int main(int argc, char **argv) {
  if (argc != 5) die("usage: ref_harness GRID NITER ALPHA_MODE OUTPREFIX");
  const char *grid = argv[1];
  int niter = atoi(argv[2]);
  int alpha_mode = atoi(argv[3]);
  std::string prefix = argv[4];

  int fake_argc = 1;
  op_init(fake_argc, argv, 0);

  FILE *fp = fopen(grid, "r");
  if (!fp) die("cannot open grid");
  int nnode, ncell, nedge, nbedge;
  if (fscanf(fp, "%d %d %d %d", &nnode, &ncell, &nedge, &nbedge) != 4)
    die("bad header");
  std::vector<int> cell(4 * ncell), edge(2 * nedge), ecell(2 * nedge),
      bedge(2 * nbedge), becell(nbedge), bound(nbedge);
  std::vector<double> x(2 * nnode), q(4 * ncell), qold(4 * ncell),
      res(4 * ncell), adt(ncell);
  for (int n = 0; n < nnode; n++)
    if (fscanf(fp, "%lf %lf", &x[2 * n], &x[2 * n + 1]) != 2) die("bad node");
  for (int n = 0; n < ncell; n++)
    if (fscanf(fp, "%d %d %d %d", &cell[4 * n], &cell[4 * n + 1],
               &cell[4 * n + 2], &cell[4 * n + 3]) != 4)
      die("bad cell");
  for (int n = 0; n < nedge; n++)
    if (fscanf(fp, "%d %d %d %d", &edge[2 * n], &edge[2 * n + 1],
               &ecell[2 * n], &ecell[2 * n + 1]) != 4)
      die("bad edge");
  for (int n = 0; n < nbedge; n++)
    if (fscanf(fp, "%d %d %d %d", &bedge[2 * n], &bedge[2 * n + 1],
               &becell[n], &bound[n]) != 4)
      die("bad bedge");
  fclose(fp);

  // constants: the same expressions as airfoil.cpp:149-164
  gam = 1.4f;
  gm1 = gam - 1.0f;
  cfl = 0.9f;
  eps = 0.05f;
  double mach_l = 0.4f;
  double alpha_l = 3.0f * atan(1.0f) / 45.0f;
  double p = 1.0f, r = 1.0f;
  double u = sqrt(gam * p / r) * mach_l;
  double e = p / (r * gm1) + 0.5f * u * u;
  qinf[0] = r;
  qinf[1] = r * u;
  qinf[2] = 0.0f;
  qinf[3] = r * e;
  if (alpha_mode == 1) {
    qinf[1] = r * u * cos(alpha_l);
    qinf[2] = r * u * sin(alpha_l);
  }
  mach = mach_l;
  alpha = alpha_l;
  for (int n = 0; n < ncell; n++)
    for (int m = 0; m < 4; m++) {
      q[4 * n + m] = qinf[m];
      res[4 * n + m] = 0.0f;
    }

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
  op_dat p_qold = op_decl_dat(cells, 4, "double", qold.data(), "p_qold");
  op_dat p_adt = op_decl_dat(cells, 1, "double", adt.data(), "p_adt");
  op_dat p_res = op_decl_dat(cells, 4, "double", res.data(), "p_res");

  FILE *hist = fopen((prefix + "_hist.txt").c_str(), "w");
  FILE *first = fopen((prefix + "_first.bin").c_str(), "wb");
  if (!hist || !first) die("cannot open outputs");
  std::vector<double> buf(4 * (size_t)ncell);

  double rms = 0, maxerr = 0, rms_k0 = 0;
  int errloc = 0;
  for (int iter = 1; iter <= niter; iter++) {
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
      if (iter == 1 && k == 0) put(first, (double *)p_adt->data, ncell);
      op_par_loop(res_calc, "res_calc", edges,
                  op_arg_dat(p_x, 0, pedge, 2, "double", OP_READ),
                  op_arg_dat(p_x, 1, pedge, 2, "double", OP_READ),
                  op_arg_dat(p_q, 0, pecell, 4, "double", OP_READ),
                  op_arg_dat(p_q, 1, pecell, 4, "double", OP_READ),
                  op_arg_dat(p_adt, 0, pecell, 1, "double", OP_READ),
                  op_arg_dat(p_adt, 1, pecell, 1, "double", OP_READ),
                  op_arg_dat(p_res, 0, pecell, 4, "double", OP_INC),
                  op_arg_dat(p_res, 1, pecell, 4, "double", OP_INC));
      if (iter == 1 && k == 0) put(first, (double *)p_res->data, 4 * (size_t)ncell);
      op_par_loop(bres_calc, "bres_calc", bedges,
                  op_arg_dat(p_x, 0, pbedge, 2, "double", OP_READ),
                  op_arg_dat(p_x, 1, pbedge, 2, "double", OP_READ),
                  op_arg_dat(p_q, 0, pbecell, 4, "double", OP_READ),
                  op_arg_dat(p_adt, 0, pbecell, 1, "double", OP_READ),
                  op_arg_dat(p_res, 0, pbecell, 4, "double", OP_INC),
                  op_arg_dat(p_bound, -1, OP_ID, 1, "int", OP_READ));
      if (iter == 1 && k == 0) put(first, (double *)p_res->data, 4 * (size_t)ncell);
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
      if (iter == 1 && k == 0) put(first, (double *)p_q->data, 4 * (size_t)ncell);
      if (k == 0) rms_k0 = rms;
    }
    double rms_raw = rms;
    rms = sqrt(rms / (double)op_get_size(cells));
    fprintf(hist, "%d %.17e %.17e %.17e %.17e %d\n", iter, rms_k0, rms_raw, rms,
            maxerr, errloc);
  }
  fclose(hist);
  fclose(first);

  op_fetch_data(p_q, buf.data());
  FILE *qf = fopen((prefix + "_q.bin").c_str(), "wb");
  if (!qf) die("cannot open q output");
  put(qf, buf.data(), 4 * (size_t)ncell);
  fclose(qf);
  op_exit();
  return 0;
}
