// The airfoil application on this library's OP2 API
// (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:72-315), as a callable entry
// point for tests and benchmarks.
#include "op2amd_airfoil.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "op2_hip_airfoil.h"
#include "op2_hip_rt.h"

namespace {

double now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Airfoil {
  const char *rt = "double";  // storage type of p_adt and p_res
  op_set nodes, edges, bedges, cells;
  op_map pedge, pecell, pbedge, pbecell, pcell;
  op_dat p_bound, p_x, p_q, p_qold, p_adt, p_res;
  double rms = 0, maxerr = 0, rms_k0 = 0;
  int errloc = 0;

  // airfoil.cpp:223-278
  void step() {
    op_par_loop_airfoil_1_save_soln("save_soln", cells,
                                    op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_READ),
                                    op_arg_dat(p_qold, -1, OP_ID, 4, "double", OP_WRITE));
    for (int k = 0; k < 2; k++) {
      op_par_loop_airfoil_2_adt_calc("adt_calc", cells,
                                     op_arg_dat(p_x, 0, pcell, 2, "double", OP_READ),
                                     op_arg_dat(p_x, 1, pcell, 2, "double", OP_READ),
                                     op_arg_dat(p_x, 2, pcell, 2, "double", OP_READ),
                                     op_arg_dat(p_x, 3, pcell, 2, "double", OP_READ),
                                     op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_READ),
                                     op_arg_dat(p_adt, -1, OP_ID, 1, rt, OP_WRITE));
      op_par_loop_airfoil_3_res_calc("res_calc", edges,
                                     op_arg_dat(p_x, 0, pedge, 2, "double", OP_READ),
                                     op_arg_dat(p_x, 1, pedge, 2, "double", OP_READ),
                                     op_arg_dat(p_q, 0, pecell, 4, "double", OP_READ),
                                     op_arg_dat(p_q, 1, pecell, 4, "double", OP_READ),
                                     op_arg_dat(p_adt, 0, pecell, 1, rt, OP_READ),
                                     op_arg_dat(p_adt, 1, pecell, 1, rt, OP_READ),
                                     op_arg_dat(p_res, 0, pecell, 4, rt, OP_INC),
                                     op_arg_dat(p_res, 1, pecell, 4, rt, OP_INC));
      op_par_loop_airfoil_4_bres_calc("bres_calc", bedges,
                                      op_arg_dat(p_x, 0, pbedge, 2, "double", OP_READ),
                                      op_arg_dat(p_x, 1, pbedge, 2, "double", OP_READ),
                                      op_arg_dat(p_q, 0, pbecell, 4, "double", OP_READ),
                                      op_arg_dat(p_adt, 0, pbecell, 1, rt, OP_READ),
                                      op_arg_dat(p_res, 0, pbecell, 4, rt, OP_INC),
                                      op_arg_dat(p_bound, -1, OP_ID, 1, "int", OP_READ));
      rms = 0.0;
      maxerr = 0.0;
      errloc = 0;
      op_par_loop_airfoil_5_update("update", cells,
                                   op_arg_dat(p_qold, -1, OP_ID, 4, "double", OP_READ),
                                   op_arg_dat(p_q, -1, OP_ID, 4, "double", OP_WRITE),
                                   op_arg_dat(p_res, -1, OP_ID, 4, rt, OP_RW),
                                   op_arg_dat(p_adt, -1, OP_ID, 1, rt, OP_READ),
                                   op_arg_gbl(&rms, 1, "double", OP_INC),
                                   op_arg_gbl(&maxerr, 1, "double", OP_MAX),
                                   op_arg_idx(-1, OP_ID),
                                   op_arg_info(&errloc, 1, "int", 5));
      if (k == 0) rms_k0 = rms;
    }
  }
};

void record(const op2amd_airfoil_result *r, int it, double rms_raw, double rms_k0,
            double maxerr, int errloc, int ncell) {
  if (r->rms) r->rms[it] = std::sqrt(rms_raw / (double)ncell);
  if (r->rms_k0) r->rms_k0[it] = rms_k0;
  if (r->maxerr) r->maxerr[it] = maxerr;
  if (r->errloc) r->errloc[it] = errloc;
}

}  // namespace

extern "C" int op2amd_airfoil_run(const op2amd_mesh *m, const op2amd_airfoil_opts *o,
                                  op2amd_airfoil_result *r) {
  if (!m || !o || !r || o->niter < 1 || o->warmup < 0 || o->warmup > o->niter)
    return 1;
  double t_setup0 = now();
  std::string soa_tok = "OP_AUTO_SOA", ren_tok = "OP_HIP_RENUMBER=pecell";
  char prog[] = "airfoil";
  char *argv[3] = {prog, nullptr, nullptr};
  int argc = 1;
  if (o->soa) argv[argc++] = &soa_tok[0];
  if (o->renumber == 2) argv[argc++] = &ren_tok[0];  // transparent renumbering
  op_init(argc, argv, 0);
  op2h_set_plan_builder(o->plan_on_device);
  op2h_enable_kernel_timing(o->kernel_timing);

  // constants (airfoil.cpp:149-164); <math.h> in C++ resolves atan(1.0f) to
  // the float overload, reproduced with std::atan(float)
  double gam = 1.4f, gm1 = gam - 1.0f, cfl = 0.9f, eps = 0.05f;
  double mach = 0.4f;
  double alpha = 3.0f * std::atan(1.0f) / 45.0f;
  double p = 1.0f, rr = 1.0f;
  double u = sqrt(gam * p / rr) * mach;
  double e = p / (rr * gm1) + 0.5f * u * u;
  double qinf[4] = {rr, rr * u, 0.0f, rr * e};
  if (o->alpha_mode == 1) {
    qinf[1] = rr * u * cos(alpha);
    qinf[2] = rr * u * sin(alpha);
  }
  std::vector<double> q(4 * (size_t)m->ncell), res(4 * (size_t)m->ncell, 0.0);
  for (int n = 0; n < m->ncell; n++)
    for (int k = 0; k < 4; k++) q[4 * (size_t)n + k] = qinf[k];

  Airfoil A;
  if (o->mixed) A.rt = "float";
  A.nodes = op_decl_set(m->nnode, "nodes");
  A.edges = op_decl_set(m->nedge, "edges");
  A.bedges = op_decl_set(m->nbedge, "bedges");
  A.cells = op_decl_set(m->ncell, "cells");
  A.pedge = op_decl_map(A.edges, A.nodes, 2, m->edge, "pedge");
  A.pecell = op_decl_map(A.edges, A.cells, 2, m->ecell, "pecell");
  A.pbedge = op_decl_map(A.bedges, A.nodes, 2, m->bedge, "pbedge");
  A.pbecell = op_decl_map(A.bedges, A.cells, 1, m->becell, "pbecell");
  A.pcell = op_decl_map(A.cells, A.nodes, 4, m->cell, "pcell");
  A.p_bound = op_decl_dat(A.bedges, 1, "int", m->bound, "p_bound");
  A.p_x = op_decl_dat(A.nodes, 2, "double", m->x, "p_x");
  A.p_q = op_decl_dat(A.cells, 4, "double", q.data(), "p_q");
  A.p_qold = op_decl_dat(A.cells, 4, "double", q.data(), "p_qold");
  std::vector<float> resf(4 * (size_t)m->ncell, 0.0f);
  if (o->mixed) {
    A.p_adt = op_decl_dat_temp(A.cells, 1, "float", (float *)nullptr, "p_adt");
    A.p_res = op_decl_dat(A.cells, 4, "float", resf.data(), "p_res");
  } else {
    A.p_adt = op_decl_dat_temp(A.cells, 1, "double", (double *)nullptr, "p_adt");
    A.p_res = op_decl_dat(A.cells, 4, "double", res.data(), "p_res");
  }
  op_decl_const2("gam", 1, "double", &gam);
  op_decl_const2("gm1", 1, "double", &gm1);
  op_decl_const2("cfl", 1, "double", &cfl);
  op_decl_const2("eps", 1, "double", &eps);
  op_decl_const2("mach", 1, "double", &mach);
  op_decl_const2("alpha", 1, "double", &alpha);
  op_decl_const2("qinf", 4, "double", qinf);
  if (o->renumber == 1) op_renumber(A.pecell);
  std::vector<int> cell_new_of_old(m->ncell);
  op2h_set_permutation(A.cells, cell_new_of_old.data());
  std::vector<int> cell_old_of_new(m->ncell);
  for (int c = 0; c < m->ncell; ++c) cell_old_of_new[cell_new_of_old[c]] = c;

  // compulsory bytes per loop call (same argument lists as step())
  {
    op_arg a1[2] = {op_arg_dat(A.p_q, -1, OP_ID, 4, "double", OP_READ),
                    op_arg_dat(A.p_qold, -1, OP_ID, 4, "double", OP_WRITE)};
    op_arg a2[6] = {op_arg_dat(A.p_x, 0, A.pcell, 2, "double", OP_READ),
                    op_arg_dat(A.p_x, 1, A.pcell, 2, "double", OP_READ),
                    op_arg_dat(A.p_x, 2, A.pcell, 2, "double", OP_READ),
                    op_arg_dat(A.p_x, 3, A.pcell, 2, "double", OP_READ),
                    op_arg_dat(A.p_q, -1, OP_ID, 4, "double", OP_READ),
                    op_arg_dat(A.p_adt, -1, OP_ID, 1, A.rt, OP_WRITE)};
    op_arg a3[8] = {op_arg_dat(A.p_x, 0, A.pedge, 2, "double", OP_READ),
                    op_arg_dat(A.p_x, 1, A.pedge, 2, "double", OP_READ),
                    op_arg_dat(A.p_q, 0, A.pecell, 4, "double", OP_READ),
                    op_arg_dat(A.p_q, 1, A.pecell, 4, "double", OP_READ),
                    op_arg_dat(A.p_adt, 0, A.pecell, 1, A.rt, OP_READ),
                    op_arg_dat(A.p_adt, 1, A.pecell, 1, A.rt, OP_READ),
                    op_arg_dat(A.p_res, 0, A.pecell, 4, A.rt, OP_INC),
                    op_arg_dat(A.p_res, 1, A.pecell, 4, A.rt, OP_INC)};
    op_arg a4[6] = {op_arg_dat(A.p_x, 0, A.pbedge, 2, "double", OP_READ),
                    op_arg_dat(A.p_x, 1, A.pbedge, 2, "double", OP_READ),
                    op_arg_dat(A.p_q, 0, A.pbecell, 4, "double", OP_READ),
                    op_arg_dat(A.p_adt, 0, A.pbecell, 1, A.rt, OP_READ),
                    op_arg_dat(A.p_res, 0, A.pbecell, 4, A.rt, OP_INC),
                    op_arg_dat(A.p_bound, -1, OP_ID, 1, "int", OP_READ)};
    op_arg a5[4] = {op_arg_dat(A.p_qold, -1, OP_ID, 4, "double", OP_READ),
                    op_arg_dat(A.p_q, -1, OP_ID, 4, "double", OP_WRITE),
                    op_arg_dat(A.p_res, -1, OP_ID, 4, A.rt, OP_RW),
                    op_arg_dat(A.p_adt, -1, OP_ID, 1, A.rt, OP_READ)};
    r->loop_bytes[0] = op2h_algorithmic_bytes(A.cells, 2, a1);
    r->loop_bytes[1] = op2h_algorithmic_bytes(A.cells, 6, a2);
    r->loop_bytes[2] = op2h_algorithmic_bytes(A.edges, 8, a3);
    r->loop_bytes[3] = op2h_algorithmic_bytes(A.bedges, 6, a4);
    r->loop_bytes[4] = op2h_algorithmic_bytes(A.cells, 4, a5);
  }

  if (o->mode == OP2AMD_MODE_PLAIN) {
    // the reference's execution model: a device sync after every loop; the
    // indirect loops on their default (LDS-staged, two-level coloured) plan
    // unless a staged or coloured variant is named
    op2h_set_sync_mode(OP2H_SYNC_EVERY_LOOP);
    op2h_set_loop_plan(nullptr, OP2H_PLAN_AUTO, 0);
    for (auto pl : {std::make_pair("res_calc", std::make_pair(o->plan_res, o->block_res)),
                    std::make_pair("bres_calc", std::make_pair(o->plan_bres, o->block_bres))})
      if (pl.second.first != OP2H_PLAN_AUTO || pl.second.second != 0) {
        if (pl.second.first != OP2H_PLAN_AUTO && pl.second.first != OP2H_PLAN_STAGED &&
            pl.second.first != OP2H_PLAN_COLOR) {
          printf("op2amd_airfoil_run: plain mode runs staged or coloured plans only\n");
          op_exit();
          return 5;
        }
        op2h_set_loop_plan(pl.first, pl.second.first, pl.second.second);
      }
  } else if (o->autotune > 0) {
    op2h_set_sync_mode(OP2H_SYNC_GBL_ONLY);
    op2h_set_loop_plan(nullptr, OP2H_PLAN_AUTO, 0);
    op2h_set_autotune(1, o->autotune);
  } else {
    op2h_set_sync_mode(OP2H_SYNC_GBL_ONLY);
    op2h_set_loop_plan("res_calc", o->plan_res, o->block_res);
    op2h_set_loop_plan("bres_calc", o->plan_bres, o->block_bres);
    op2h_set_loop_plan("save_soln", OP2H_PLAN_AUTO, o->block_direct);
    op2h_set_loop_plan("update", OP2H_PLAN_AUTO, o->block_direct);
    op2h_set_loop_plan("adt_calc", OP2H_PLAN_AUTO, o->block_adt);
  }

  double t0 = 0.0;
  const int ncell = m->ncell;
  if (o->mode != OP2AMD_MODE_GRAPHED) {
    for (int it = 1; it <= o->niter; ++it) {
      if (it == o->warmup + 1) {
        op2h_device_sync();
        if (it == 1) r->setup_seconds = now() - t_setup0;
        op2h_reset_stats();
        t0 = now();
      }
      A.step();
      record(r, it - 1, A.rms, A.rms_k0, A.maxerr, A.errloc, ncell);
      if (it == o->warmup) r->setup_seconds = now() - t_setup0;
    }
    op2h_device_sync();
  } else {
    // iteration 1 eagerly: builds every plan and sizes every buffer
    if (o->warmup == 0) {
      op2h_device_sync();
      op2h_reset_stats();
      t0 = now();
    }
    A.step();
    record(r, 0, A.rms, A.rms_k0, A.maxerr, A.errloc, ncell);
    int it0 = 2;  // first iteration replayed from the graph
    while (o->autotune > 0 && !op2h_tuning_done() && it0 <= o->niter) {
      if (it0 == o->warmup + 1) {
        op2h_device_sync();
        r->setup_seconds = now() - t_setup0;
        op2h_reset_stats();
        t0 = now();
      }
      A.step();
      record(r, it0 - 1, A.rms, A.rms_k0, A.maxerr, A.errloc, ncell);
      ++it0;
    }
    // G time steps per captured graph (fewer graph launches, so fewer gaps
    // between graphs on the device), G a divisor of the timed iteration count
    // so the timed region starts and ends on graph boundaries; the replayed
    // iterations before it that do not fill a graph run eagerly
    const int eager0 = it0 - 1;
    const int gmax = std::max(1, std::min(o->graph_steps, 16));
    auto divisor_upto = [](int n, int g) {
      if (n <= 0) return 1;
      while (g > 1 && n % g) --g;
      return g;
    };
    int G, pre_launches = 0;
    if (o->warmup >= eager0) {
      G = divisor_upto(o->niter - o->warmup, gmax);
      const int pre = o->warmup - eager0;
      for (int k = 0; k < pre % G; ++k, ++it0) {
        A.step();
        record(r, it0 - 1, A.rms, A.rms_k0, A.maxerr, A.errloc, ncell);
      }
      pre_launches = pre / G;
    } else {
      G = divisor_upto(o->niter - eager0, gmax);
    }
    r->eager_iterations = it0 - 1;
    r->graph_steps = G;
    if (op2h_graph_begin() != 0) return 2;
    for (int k = 0; k < G; ++k) A.step();
    if (op2h_graph_end() != 0) return 3;
    const int launches = (o->niter - it0 + 1) / G;
    if (o->warmup >= eager0) {
      op2h_graph_launch(pre_launches);
      op2h_device_sync();
      r->setup_seconds = now() - t_setup0;
      op2h_reset_stats();
      t0 = now();
      op2h_graph_launch(launches - pre_launches);
    } else {
      op2h_graph_launch(launches);
    }
    op2h_device_sync();
    double t1 = now();
    // per-iteration gbl results from the device history; one record per
    // launch, per step [rms_k0][max_k0][loc_k0][rms_k1][max_k1][loc_k1]
    int nrec = 0;
    const int words = launches > 0 ? op2h_graph_history(0, nullptr, &nrec) : 6 * G;
    std::vector<double> hist(std::max(launches, 1) * 6 * (size_t)G);
    if (words == 6 * G && launches > 0) op2h_graph_history(launches, hist.data(), &nrec);
    if (launches > 0 && (words != 6 * G || nrec != launches)) {
      printf("op2amd_airfoil_run: graph history has %d records of %d words\n",
             nrec, words);
      op_exit();
      return 4;
    }
    for (int i = 0; i < nrec * G; ++i) {
      const double *h = &hist[(size_t)i * 6];
      double mx = h[4] > 0.0 ? h[4] : 0.0;
      int loc = 0;
      if (h[4] > 0.0) std::memcpy(&loc, &h[5], sizeof(int));
      record(r, it0 - 1 + i, 0.0 + h[3], 0.0 + h[0], mx, loc, ncell);
    }
    op2h_graph_sync();
    r->seconds = t1 - t0;
  }
  if (o->mode != OP2AMD_MODE_GRAPHED) r->seconds = now() - t0;
  const char *names[5] = {"save_soln", "adt_calc", "res_calc", "bres_calc", "update"};
  for (int l = 0; l < 5; ++l) {
    r->chosen_plan[l] = r->chosen_block[l] = -1;
    op2h_tuning_choice(names[l], &r->chosen_plan[l], &r->chosen_block[l]);
    r->loop_calls[l] = 0;
    r->loop_kernel_ms[l] = 0.0;
    op2h_loop_stats(names[l], &r->loop_calls[l], &r->loop_kernel_ms[l]);
  }
  if (r->q_out) {  // back to the input numbering
    std::vector<double> qn(4 * (size_t)m->ncell);
    op_fetch_data(A.p_q, qn.data());
    for (int c = 0; c < m->ncell; ++c)
      std::memcpy(r->q_out + 4 * (size_t)c, &qn[4 * (size_t)cell_new_of_old[c]],
                  4 * sizeof(double));
  }
  if (o->dump) {
    const std::string base(o->dump);
    op_print_dat_to_txtfile(A.p_q, (base + ".dat").c_str());
    op_print_dat_to_binfile(A.p_q, (base + ".bin").c_str());
    if (o->renumber == 1) {
      FILE *f = std::fopen((base + ".perm").c_str(), "w");
      if (!f) {
        printf("can't open file %s.perm\n", o->dump);
        exit(2);
      }
      std::fprintf(f, "%d\n", m->ncell);
      for (int c = 0; c < m->ncell; ++c) std::fprintf(f, "%d\n", cell_new_of_old[c]);
      std::fclose(f);
    }
  }
  if (r->errloc && o->renumber == 1)
    for (int it = 0; it < o->niter; ++it)
      if (r->errloc[it] >= 0 && r->errloc[it] < m->ncell)
        r->errloc[it] = cell_old_of_new[r->errloc[it]];
  op2h_graph_destroy();
  op_exit();
  return 0;
}
