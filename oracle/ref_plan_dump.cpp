// Test-infrastructure harness (oracle only; never shipped in the product).
//
// Calls the REFERENCE's own plan builder op_plan_core(..., OP_STAGE_ALL)
// (op2/src/core/op_rt_support.cpp:301-1126) for the three indirect airfoil
// loops, with the argument lists of airfoil.cpp:235-261, on a mesh in the
// new_grid.dat format, and writes the plan arrays as raw int32 records so that
// the oracle's restatement of the plan builder can be compared with it.
//
// Usage: ref_plan_dump GRID PART_SIZE OUT
// OUT layout, per loop (adt_calc, res_calc, bres_calc), all int32:
//   set_size nblocks ncolors bsize
//   offset[nblocks] nelems[nblocks] nthrcol[nblocks] thrcol[set_size]
//   ncolblk[ncolors] blkmap[nblocks]
//   ninds nindirect[ninds] ind_sizes[nblocks*ninds] ind_offs[nblocks*ninds]
//   for each ind: ind_map[nindirect]
//   for each arg with a staged map: loc_map[set_size] (as int32)
#include <stdio.h>
#include <stdlib.h>
#include <vector>

double gam, gm1, cfl, eps, mach, alpha, qinf[4];

#include "op_seq.h"

static void w(FILE *f, const int *p, size_t n) {
  if (n && fwrite(p, sizeof(int), n, f) != n) {
    fprintf(stderr, "write failed\n");
    exit(1);
  }
}

static void dump(FILE *f, op_plan *p, int set_size, int nargs) {
  int hdr[4] = {set_size, p->nblocks, p->ncolors, 0};
  hdr[3] = p->nblocks ? p->nelems[0] : 0;
  w(f, hdr, 4);
  w(f, p->offset, p->nblocks);
  w(f, p->nelems, p->nblocks);
  w(f, p->nthrcol, p->nblocks);
  w(f, p->thrcol, set_size);
  w(f, p->ncolblk, p->ncolors);
  w(f, p->blkmap, p->nblocks);
  w(f, &p->ninds_staged, 1);
  w(f, p->nindirect, p->ninds);
  w(f, p->ind_sizes, (size_t)p->nblocks * p->ninds_staged);
  w(f, p->ind_offs, (size_t)p->nblocks * p->ninds_staged);
  for (int i = 0; i < p->ninds_staged; ++i) w(f, p->ind_maps[i], p->nindirect[i]);
  std::vector<int> tmp(set_size);
  for (int m = 0; m < nargs; ++m) {
    if (!p->loc_maps[m]) continue;
    for (int e = 0; e < set_size; ++e) tmp[e] = p->loc_maps[m][e];
    w(f, tmp.data(), set_size);
  }
}

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: ref_plan_dump GRID PART_SIZE OUT\n");
    return 1;
  }
  int part_size = atoi(argv[2]);
  int one = 1;
  op_init(one, argv, 0);
  FILE *fp = fopen(argv[1], "r");
  if (!fp) return 1;
  int nnode, ncell, nedge, nbedge;
  if (fscanf(fp, "%d %d %d %d", &nnode, &ncell, &nedge, &nbedge) != 4) return 1;
  std::vector<int> cell(4 * ncell), edge(2 * nedge), ecell(2 * nedge),
      bedge(2 * nbedge), becell(nbedge), bound(nbedge);
  std::vector<double> x(2 * nnode), zeros4(4 * ncell, 0.0), zeros1(ncell, 0.0);
  for (int n = 0; n < nnode; n++)
    if (fscanf(fp, "%lf %lf", &x[2 * n], &x[2 * n + 1]) != 2) return 1;
  for (int n = 0; n < ncell; n++)
    if (fscanf(fp, "%d %d %d %d", &cell[4 * n], &cell[4 * n + 1], &cell[4 * n + 2],
               &cell[4 * n + 3]) != 4)
      return 1;
  for (int n = 0; n < nedge; n++)
    if (fscanf(fp, "%d %d %d %d", &edge[2 * n], &edge[2 * n + 1], &ecell[2 * n],
               &ecell[2 * n + 1]) != 4)
      return 1;
  for (int n = 0; n < nbedge; n++)
    if (fscanf(fp, "%d %d %d %d", &bedge[2 * n], &bedge[2 * n + 1], &becell[n],
               &bound[n]) != 4)
      return 1;
  fclose(fp);

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
  op_dat p_q = op_decl_dat(cells, 4, "double", zeros4.data(), "p_q");
  op_dat p_adt = op_decl_dat(cells, 1, "double", zeros1.data(), "p_adt");
  op_dat p_res = op_decl_dat(cells, 4, "double", zeros4.data(), "p_res");

  FILE *out = fopen(argv[3], "wb");
  if (!out) return 1;
  {
    op_arg a[6] = {op_arg_dat(p_x, 0, pcell, 2, "double", OP_READ),
                   op_arg_dat(p_x, 1, pcell, 2, "double", OP_READ),
                   op_arg_dat(p_x, 2, pcell, 2, "double", OP_READ),
                   op_arg_dat(p_x, 3, pcell, 2, "double", OP_READ),
                   op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_READ),
                   op_arg_dat(p_adt, -1, OP_ID, 1, "double", OP_WRITE)};
    int inds[6] = {0, 0, 0, 0, -1, -1};
    dump(out, op_plan_core("adt_calc", cells, part_size, 6, a, 1, inds, OP_STAGE_ALL),
         ncell, 6);
  }
  {
    op_arg a[8] = {op_arg_dat(p_x, 0, pedge, 2, "double", OP_READ),
                   op_arg_dat(p_x, 1, pedge, 2, "double", OP_READ),
                   op_arg_dat(p_q, 0, pecell, 4, "double", OP_READ),
                   op_arg_dat(p_q, 1, pecell, 4, "double", OP_READ),
                   op_arg_dat(p_adt, 0, pecell, 1, "double", OP_READ),
                   op_arg_dat(p_adt, 1, pecell, 1, "double", OP_READ),
                   op_arg_dat(p_res, 0, pecell, 4, "double", OP_INC),
                   op_arg_dat(p_res, 1, pecell, 4, "double", OP_INC)};
    int inds[8] = {0, 0, 1, 1, 2, 2, 3, 3};
    dump(out, op_plan_core("res_calc", edges, part_size, 8, a, 4, inds, OP_STAGE_ALL),
         nedge, 8);
  }
  {
    op_arg a[6] = {op_arg_dat(p_x, 0, pbedge, 2, "double", OP_READ),
                   op_arg_dat(p_x, 1, pbedge, 2, "double", OP_READ),
                   op_arg_dat(p_q, 0, pbecell, 4, "double", OP_READ),
                   op_arg_dat(p_adt, 0, pbecell, 1, "double", OP_READ),
                   op_arg_dat(p_res, 0, pbecell, 4, "double", OP_INC),
                   op_arg_dat(p_bound, -1, OP_ID, 1, "int", OP_READ)};
    int inds[6] = {0, 0, 1, 2, 3, -1};
    dump(out, op_plan_core("bres_calc", bedges, part_size, 6, a, 4, inds, OP_STAGE_ALL),
         nbedge, 6);
  }
  fclose(out);
  op_exit();
  return 0;
}
