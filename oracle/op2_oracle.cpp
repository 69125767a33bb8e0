// ORACLE — test infrastructure only (see op2_oracle.h).  CPU restatement of
// the reference airfoil path: kernels, sequential and OpenMP executors, and
// op_plan_core(OP_STAGE_ALL).  Compiled with -ffp-contract=off so that every
// multiply and add rounds separately, as the reference's g++ -O3 build without
// -march=native does.
#include "op2_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <strings.h>
#include <vector>

namespace {

// ---- kernels: expression trees as in apps/c/airfoil/airfoil_plain/dp/*.h ----
// (left-to-right evaluation of each expression is kept; float literals of the
// reference such as 0.5f and 1.0f are exact in double.)

inline void k_save_soln(const double *q, double *qold) {  // save_soln.h:1-4
  for (int n = 0; n < 4; n++) qold[n] = q[n];
}

inline void k_adt_calc(const orc_consts &c, const double *x1, const double *x2,
                       const double *x3, const double *x4, const double *q,
                       double *adt) {  // adt_calc.h:1-27
  double ri = 1.0 / q[0];
  double u = ri * q[1];
  double v = ri * q[2];
  double cs = std::sqrt(c.gam * c.gm1 * (ri * q[3] - 0.5 * (u * u + v * v)));
  const double *a[5] = {x1, x2, x3, x4, x1};
  double acc = 0.0;
  for (int s = 0; s < 4; ++s) {
    double dx = a[s + 1][0] - a[s][0];
    double dy = a[s + 1][1] - a[s][1];
    double t = std::fabs(u * dy - v * dx) + cs * std::sqrt(dx * dx + dy * dy);
    acc = (s == 0) ? t : acc + t;
  }
  *adt = acc / c.cfl;
}

inline void k_res_calc(const orc_consts &c, const double *x1, const double *x2,
                       const double *q1, const double *q2, const double *adt1,
                       const double *adt2, double *res1,
                       double *res2) {  // res_calc.h:1-33
  double dx = x1[0] - x2[0];
  double dy = x1[1] - x2[1];
  double ri = 1.0 / q1[0];
  double p1 = c.gm1 * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  double vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0 / q2[0];
  double p2 = c.gm1 * (q2[3] - 0.5 * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  double vol2 = ri * (q2[1] * dy - q2[2] * dx);
  double mu = 0.5 * ((*adt1) + (*adt2)) * c.eps;
  double f;
  f = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  res1[0] += f;
  res2[0] -= f;
  f = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) +
      mu * (q1[1] - q2[1]);
  res1[1] += f;
  res2[1] -= f;
  f = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) +
      mu * (q1[2] - q2[2]);
  res1[2] += f;
  res2[2] -= f;
  f = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
  res1[3] += f;
  res2[3] -= f;
}

inline void k_bres_calc(const orc_consts &c, const double *x1, const double *x2,
                        const double *q1, const double *adt1, double *res1,
                        const int *bound) {  // bres_calc.h:1-35
  double dx = x1[0] - x2[0];
  double dy = x1[1] - x2[1];
  double ri = 1.0 / q1[0];
  double p1 = c.gm1 * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  if (*bound == 1) {
    res1[1] += +p1 * dy;
    res1[2] += -p1 * dx;
    return;
  }
  const double *qi = c.qinf;
  double vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0 / qi[0];
  double p2 = c.gm1 * (qi[3] - 0.5 * ri * (qi[1] * qi[1] + qi[2] * qi[2]));
  double vol2 = ri * (qi[1] * dy - qi[2] * dx);
  double mu = (*adt1) * c.eps;
  double f;
  f = 0.5 * (vol1 * q1[0] + vol2 * qi[0]) + mu * (q1[0] - qi[0]);
  res1[0] += f;
  f = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * qi[1] + p2 * dy) +
      mu * (q1[1] - qi[1]);
  res1[1] += f;
  f = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * qi[2] - p2 * dx) +
      mu * (q1[2] - qi[2]);
  res1[2] += f;
  f = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (qi[3] + p2)) + mu * (q1[3] - qi[3]);
  res1[3] += f;
}

inline void k_update(const double *qold, double *q, double *res,
                     const double *adt, double *rms, double *maxerr,
                     const int *idx, int *errloc) {  // update.h:1-19
  double adti = 1.0 / (*adt);
  for (int n = 0; n < 4; n++) {
    double del = adti * res[n];
    q[n] = qold[n] - del;
    res[n] = 0.0;
    double sqdel = del * del;
    *rms += sqdel;
    if (sqdel > *maxerr) {
      *maxerr = sqdel;
      *errloc = *idx;
    }
  }
}

// ---- element-range bodies shared by the executors ----

void adt_range(const op2amd_mesh *m, const orc_consts &c, const double *q,
               double *adt, int lo, int hi) {
  for (int n = lo; n < hi; ++n) {
    const int *cn = m->cell + 4 * (size_t)n;
    k_adt_calc(c, m->x + 2 * (size_t)cn[0], m->x + 2 * (size_t)cn[1],
               m->x + 2 * (size_t)cn[2], m->x + 2 * (size_t)cn[3],
               q + 4 * (size_t)n, adt + n);
  }
}

void res_range(const op2amd_mesh *m, const orc_consts &c, const double *q,
               const double *adt, double *res, int lo, int hi) {
  for (int e = lo; e < hi; ++e) {
    const int *en = m->edge + 2 * (size_t)e, *ec = m->ecell + 2 * (size_t)e;
    k_res_calc(c, m->x + 2 * (size_t)en[0], m->x + 2 * (size_t)en[1],
               q + 4 * (size_t)ec[0], q + 4 * (size_t)ec[1], adt + ec[0],
               adt + ec[1], res + 4 * (size_t)ec[0], res + 4 * (size_t)ec[1]);
  }
}

void bres_range(const op2amd_mesh *m, const orc_consts &c, const double *q,
                const double *adt, double *res, int lo, int hi) {
  for (int e = lo; e < hi; ++e) {
    const int *bn = m->bedge + 2 * (size_t)e;
    int cc = m->becell[e];
    k_bres_calc(c, m->x + 2 * (size_t)bn[0], m->x + 2 * (size_t)bn[1],
                q + 4 * (size_t)cc, adt + cc, res + 4 * (size_t)cc,
                m->bound + e);
  }
}

void update_range(const double *qold, double *q, double *res,
                  const double *adt, double *rms, double *maxerr, int *errloc,
                  int lo, int hi) {
  for (int n = lo; n < hi; ++n)
    k_update(qold + 4 * (size_t)n, q + 4 * (size_t)n, res + 4 * (size_t)n,
             adt + n, rms, maxerr, &n, errloc);
}

double now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" void orc_airfoil_consts(int alpha_mode, orc_consts *c) {
  // airfoil.cpp:149-164; float literals assigned to doubles as written there.
  c->gam = 1.4f;
  c->gm1 = c->gam - 1.0f;
  c->cfl = 0.9f;
  c->eps = 0.05f;
  c->mach = 0.4f;
  // <math.h> in C++ resolves atan(1.0f) to the float overload: float arithmetic.
  c->alpha = 3.0f * std::atan(1.0f) / 45.0f;
  double p = 1.0f, r = 1.0f;
  double u = sqrt(c->gam * p / r) * c->mach;
  double e = p / (r * c->gm1) + 0.5f * u * u;
  c->qinf[0] = r;
  c->qinf[1] = r * u;
  c->qinf[2] = 0.0f;
  c->qinf[3] = r * e;
  if (alpha_mode == 1) {
    c->qinf[1] = r * u * cos(c->alpha);
    c->qinf[2] = r * u * sin(c->alpha);
  }
}

extern "C" void orc_save_soln(int ncell, const double *q, double *qold) {
  for (int n = 0; n < ncell; ++n) k_save_soln(q + 4 * (size_t)n, qold + 4 * (size_t)n);
}
extern "C" void orc_adt_calc(const op2amd_mesh *m, const orc_consts *c,
                             const double *q, double *adt) {
  adt_range(m, *c, q, adt, 0, m->ncell);
}
extern "C" void orc_res_calc(const op2amd_mesh *m, const orc_consts *c,
                             const double *q, const double *adt, double *res) {
  res_range(m, *c, q, adt, res, 0, m->nedge);
}
extern "C" void orc_bres_calc(const op2amd_mesh *m, const orc_consts *c,
                              const double *q, const double *adt, double *res) {
  bres_range(m, *c, q, adt, res, 0, m->nbedge);
}
extern "C" void orc_update(int ncell, const double *qold, double *q,
                           double *res, const double *adt, double *rms,
                           double *maxerr, int *errloc) {
  update_range(qold, q, res, adt, rms, maxerr, errloc, 0, ncell);
}

// ---------------------------------------------------------------------------
// op_plan_core(OP_STAGE_ALL), one rank: core_size = size, no exec halo.
// ---------------------------------------------------------------------------
extern "C" int orc_plan_build(int set_size, int nargs, const int *const *maps,
                              const int *mdim, const int *idx, const int *inds,
                              const int *acc, const int *dat_bytes,
                              const int *to_size, int part_size,
                              orc_plan *P) {
  std::memset(P, 0, sizeof(*P));
  int ninds = 0;
  for (int m = 0; m < nargs; ++m) ninds = std::max(ninds, inds[m] + 1);
  // op_rt_support.cpp:355-378: worst-case staged bytes and block size
  int maxbytes = 0;
  for (int m = 0; m < nargs; ++m)
    if (inds[m] >= 0) maxbytes += dat_bytes[m];
  int bsize = part_size;
  if (bsize == 0 && maxbytes > 0)
    bsize = std::max((24 * 1024 / (64 * maxbytes)) * 64, 256);
  else if (bsize == 0)
    bsize = 256;
  int nblocks = set_size > 0 ? (set_size + bsize - 1) / bsize : 0;
  P->bsize = bsize;
  P->nblocks = nblocks;
  P->ninds = ninds;
  P->offset = (int *)std::calloc(std::max(nblocks, 1), sizeof(int));
  P->nelems = (int *)std::calloc(std::max(nblocks, 1), sizeof(int));
  P->nthrcol = (int *)std::calloc(std::max(nblocks, 1), sizeof(int));
  P->blk_col = (int *)std::calloc(std::max(nblocks, 1), sizeof(int));
  P->blkmap = (int *)std::calloc(std::max(nblocks, 1), sizeof(int));
  P->thrcol = (int *)std::calloc(std::max(set_size, 1), sizeof(int));
  P->ind_sizes = (int *)std::calloc(std::max(nblocks * ninds, 1), sizeof(int));
  P->ind_offs = (int *)std::calloc(std::max(nblocks * ninds, 1), sizeof(int));
  P->nindirect = (int *)std::calloc(std::max(ninds, 1), sizeof(int));
  P->loc_map = (short *)std::calloc(std::max((size_t)nargs * set_size, (size_t)1),
                                    sizeof(short));
  // ind_map: for each ind, (number of args using it) * set_size entries max
  std::vector<int> nuse(ninds, 0);
  for (int m = 0; m < nargs; ++m)
    if (inds[m] >= 0) nuse[inds[m]]++;
  std::vector<size_t> ind_base(ninds + 1, 0);
  for (int i = 0; i < ninds; ++i)
    ind_base[i + 1] = ind_base[i] + (size_t)nuse[i] * set_size;
  P->ind_map = (int *)std::calloc(std::max(ind_base[ninds], (size_t)1), sizeof(int));

  auto mapv = [&](int m, int e) { return maps[m][idx[m] + (size_t)e * mdim[m]]; };
  auto conflicting = [&](int m) { return inds[m] >= 0 && (acc[m] == 3 || acc[m] == 2); };

  std::vector<std::vector<unsigned>> work(ninds);
  for (int i = 0; i < ninds; ++i) {
    int m2 = 0;
    while (inds[m2] != i) ++m2;
    work[i].assign(std::max(to_size[m2], 1), 0u);
  }
  std::vector<int> work2;
  for (int b = 0; b < nblocks; ++b) {
    int lo = b * bsize, hi = std::min(lo + bsize, set_size);
    P->offset[b] = lo;
    P->nelems[b] = hi - lo;
    // op_rt_support.cpp:609-682: unique sorted indirect elements per block
    for (int i = 0; i < ninds; ++i) {
      work2.clear();
      for (int m = 0; m < nargs; ++m)
        if (inds[m] == i)
          for (int e = lo; e < hi; ++e) work2.push_back(mapv(m, e));
      std::sort(work2.begin(), work2.end());
      work2.erase(std::unique(work2.begin(), work2.end()), work2.end());
      int *dst = P->ind_map + ind_base[i];
      for (size_t k = 0; k < work2.size(); ++k) {
        dst[P->nindirect[i]++] = work2[k];
        work[i][work2[k]] = (unsigned)k;
      }
      for (int m = 0; m < nargs; ++m)
        if (inds[m] == i)
          for (int e = lo; e < hi; ++e)
            P->loc_map[(size_t)m * set_size + e] = (short)work[i][mapv(m, e)];
      P->ind_offs[i + b * ninds] =
          b == 0 ? 0 : P->ind_offs[i + (b - 1) * ninds] + P->ind_sizes[i + (b - 1) * ninds];
      P->ind_sizes[i + b * ninds] = P->nindirect[i] - P->ind_offs[i + b * ninds];
    }
    // op_rt_support.cpp:684-754: greedy thread colouring, 32 colours per pass
    for (int e = lo; e < hi; ++e) P->thrcol[e] = -1;
    int repeat = 1, ncolor = 0, ncolors = 0;
    while (repeat) {
      repeat = 0;
      for (int m = 0; m < nargs; ++m)
        if (inds[m] >= 0)
          for (int e = lo; e < hi; ++e) work[inds[m]][mapv(m, e)] = 0;
      for (int e = lo; e < hi; ++e) {
        if (P->thrcol[e] != -1) continue;
        unsigned mask = 0;
        for (int m = 0; m < nargs; ++m)
          if (conflicting(m)) mask |= work[inds[m]][mapv(m, e)];
        int color = ffs((int)~mask) - 1;
        if (color == -1) {
          repeat = 1;
        } else {
          P->thrcol[e] = ncolor + color;
          ncolors = std::max(ncolors, ncolor + color + 1);
          for (int m = 0; m < nargs; ++m)
            if (conflicting(m)) work[inds[m]][mapv(m, e)] |= 1u << color;
        }
      }
      ncolor += 32;
    }
    P->nthrcol[b] = ncolors;
  }
  // op_rt_support.cpp:798-884: greedy block colouring
  for (int b = 0; b < nblocks; ++b) P->blk_col[b] = -1;
  int repeat = 1, ncolor = 0, ncolors = 0;
  while (repeat) {
    repeat = 0;
    for (int i = 0; i < ninds; ++i) std::fill(work[i].begin(), work[i].end(), 0u);
    for (int b = 0; b < nblocks; ++b) {
      if (P->blk_col[b] != -1) continue;
      int lo = P->offset[b], hi = lo + P->nelems[b];
      unsigned mask = 0;
      for (int m = 0; m < nargs; ++m)
        if (conflicting(m))
          for (int e = lo; e < hi; ++e) mask |= work[inds[m]][mapv(m, e)];
      int color = ffs((int)~mask) - 1;
      if (color == -1) {
        repeat = 1;
      } else {
        P->blk_col[b] = ncolor + color;
        ncolors = std::max(ncolors, ncolor + color + 1);
        for (int m = 0; m < nargs; ++m)
          if (conflicting(m))
            for (int e = lo; e < hi; ++e) work[inds[m]][mapv(m, e)] |= 1u << color;
      }
    }
    ncolor += 32;
  }
  P->ncolors = ncolors;
  // op_rt_support.cpp:901-922: blocks grouped by colour, ascending index
  P->ncolblk = (int *)std::calloc(std::max(ncolors, 1), sizeof(int));
  for (int b = 0; b < nblocks; ++b) P->ncolblk[P->blk_col[b]]++;
  std::vector<int> start(ncolors + 1, 0);
  for (int c = 0; c < ncolors; ++c) start[c + 1] = start[c] + P->ncolblk[c];
  for (int b = 0; b < nblocks; ++b) P->blkmap[start[P->blk_col[b]]++] = b;
  return 0;
}

extern "C" void orc_plan_free(orc_plan *P) {
  std::free(P->offset);
  std::free(P->nelems);
  std::free(P->nthrcol);
  std::free(P->thrcol);
  std::free(P->blk_col);
  std::free(P->ncolblk);
  std::free(P->blkmap);
  std::free(P->ind_sizes);
  std::free(P->ind_offs);
  std::free(P->ind_map);
  std::free(P->nindirect);
  std::free(P->loc_map);
  std::memset(P, 0, sizeof(*P));
}

// ---------------------------------------------------------------------------
// Full airfoil run, sequential or OpenMP.
// ---------------------------------------------------------------------------
namespace {

struct OmpLoopPlan {
  orc_plan p{};
  ~OmpLoopPlan() { orc_plan_free(&p); }
};

// Run an indirect loop body colour by colour (openmp/loop_host.hpp.jinja:246-285).
template <class Body>
void run_blocks(const orc_plan &p, int nthreads, Body body) {
  int block_offset = 0;
  for (int col = 0; col < p.ncolors; ++col) {
    int nb = p.ncolblk[col];
#pragma omp parallel for num_threads(nthreads)
    for (int k = 0; k < nb; ++k) {
      int b = p.blkmap[block_offset + k];
      body(p.offset[b], p.offset[b] + p.nelems[b]);
    }
    block_offset += nb;
  }
}

}  // namespace

static int run_with_consts(const op2amd_mesh *m, int niter, const orc_consts &c,
                           int exec, int nthreads, int part_size,
                           double *rms_k0_raw, double *rms_out, double *maxerr_out,
                           int *errloc_out, double *q_out, double *loop_seconds);

extern "C" int orc_airfoil_run(const op2amd_mesh *m, int niter, int alpha_mode,
                               int exec, int nthreads, int part_size,
                               double *rms_k0_raw, double *rms_out,
                               double *maxerr_out, int *errloc_out,
                               double *q_out, double *loop_seconds) {
  orc_consts c;
  orc_airfoil_consts(alpha_mode, &c);
  return run_with_consts(m, niter, c, exec, nthreads, part_size, rms_k0_raw, rms_out,
                         maxerr_out, errloc_out, q_out, loop_seconds);
}

extern "C" int orc_airfoil_run_qinf(const op2amd_mesh *m, int niter,
                                    const double *qinf, int exec, int nthreads,
                                    int part_size, double *rms_k0_raw,
                                    double *rms_out, double *maxerr_out,
                                    int *errloc_out, double *q_out,
                                    double *loop_seconds) {
  orc_consts c;
  orc_airfoil_consts(0, &c);
  for (int k = 0; k < 4; ++k) c.qinf[k] = qinf[k];
  return run_with_consts(m, niter, c, exec, nthreads, part_size, rms_k0_raw, rms_out,
                         maxerr_out, errloc_out, q_out, loop_seconds);
}

static int run_with_consts(const op2amd_mesh *m, int niter, const orc_consts &c,
                           int exec, int nthreads, int part_size,
                           double *rms_k0_raw, double *rms_out, double *maxerr_out,
                           int *errloc_out, double *q_out, double *loop_seconds) {
  const int nc = m->ncell;
  std::vector<double> q(4 * (size_t)nc), qold(4 * (size_t)nc),
      res(4 * (size_t)nc, 0.0), adt(nc);
  for (int n = 0; n < nc; ++n)
    for (int k = 0; k < 4; ++k) q[4 * (size_t)n + k] = c.qinf[k];
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  double tl[5] = {0, 0, 0, 0, 0};

  OmpLoopPlan adt_plan, res_plan, bres_plan;
  if (exec == 1) {
    // argument descriptors as declared at airfoil.cpp:235-261
    const int *adt_maps[6] = {m->cell, m->cell, m->cell, m->cell, nullptr, nullptr};
    int adt_mdim[6] = {4, 4, 4, 4, 0, 0}, adt_idx[6] = {0, 1, 2, 3, -1, -1};
    int adt_inds[6] = {0, 0, 0, 0, -1, -1}, adt_acc[6] = {0, 0, 0, 0, 0, 1};
    int adt_bytes[6] = {16, 16, 16, 16, 32, 8};
    int adt_to[6] = {m->nnode, m->nnode, m->nnode, m->nnode, nc, nc};
    orc_plan_build(nc, 6, adt_maps, adt_mdim, adt_idx, adt_inds, adt_acc,
                   adt_bytes, adt_to, part_size, &adt_plan.p);
    const int *res_maps[8] = {m->edge, m->edge, m->ecell, m->ecell,
                              m->ecell, m->ecell, m->ecell, m->ecell};
    int res_mdim[8] = {2, 2, 2, 2, 2, 2, 2, 2}, res_idx[8] = {0, 1, 0, 1, 0, 1, 0, 1};
    int res_inds[8] = {0, 0, 1, 1, 2, 2, 3, 3}, res_acc[8] = {0, 0, 0, 0, 0, 0, 3, 3};
    int res_bytes[8] = {16, 16, 32, 32, 8, 8, 32, 32};
    int res_to[8] = {m->nnode, m->nnode, nc, nc, nc, nc, nc, nc};
    orc_plan_build(m->nedge, 8, res_maps, res_mdim, res_idx, res_inds, res_acc,
                   res_bytes, res_to, part_size, &res_plan.p);
    const int *bres_maps[6] = {m->bedge, m->bedge, m->becell, m->becell, m->becell, nullptr};
    int bres_mdim[6] = {2, 2, 1, 1, 1, 0}, bres_idx[6] = {0, 1, 0, 0, 0, -1};
    int bres_inds[6] = {0, 0, 1, 2, 3, -1}, bres_acc[6] = {0, 0, 0, 0, 3, 0};
    int bres_bytes[6] = {16, 16, 32, 8, 32, 4};
    int bres_to[6] = {m->nnode, m->nnode, nc, nc, nc, m->nbedge};
    orc_plan_build(m->nbedge, 6, bres_maps, bres_mdim, bres_idx, bres_inds,
                   bres_acc, bres_bytes, bres_to, part_size, &bres_plan.p);
  }

  // exec 2: the sequential executor's results, evaluated in parallel.  Each
  // interior edge's two increments are evaluated once into a scratch pair
  // started at -0.0 (so it holds exactly +f and -f: -0.0 + f == f and
  // -0.0 - f == -f for every f, signed zeros included); each cell then adds its
  // increments in the sequential order (interior edges ascending, then
  // boundary edges ascending), and x - f == x + (-f) exactly.  The update's
  // per-component squares are summed and compared serially in element order.
  std::vector<int> inc_off, inc_ent;  // per cell: 2*edge+side, then bedge ids
  std::vector<double> eflux, bflux, sq;
  if (exec == 2) {
    inc_off.assign(nc + 1, 0);
    for (size_t i = 0; i < 2 * (size_t)m->nedge; ++i) inc_off[m->ecell[i] + 1]++;
    for (int b = 0; b < m->nbedge; ++b) inc_off[m->becell[b] + 1]++;
    for (int n = 0; n < nc; ++n) inc_off[n + 1] += inc_off[n];
    inc_ent.resize(inc_off[nc]);
    std::vector<int> fill(inc_off.begin(), inc_off.end() - 1);
    for (size_t i = 0; i < 2 * (size_t)m->nedge; ++i) inc_ent[fill[m->ecell[i]]++] = (int)i;
    for (int b = 0; b < m->nbedge; ++b)
      inc_ent[fill[m->becell[b]]++] = -1 - b;  // after every interior edge
    eflux.resize(8 * (size_t)m->nedge);
    bflux.resize(4 * (size_t)m->nbedge);
    sq.resize(4 * (size_t)nc);
  }
  auto res_bres_ordered = [&]() {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int e = 0; e < m->nedge; ++e) {
      double *f = &eflux[8 * (size_t)e];
      for (int k = 0; k < 8; ++k) f[k] = -0.0;
      const int *en = m->edge + 2 * (size_t)e, *ec = m->ecell + 2 * (size_t)e;
      k_res_calc(c, m->x + 2 * (size_t)en[0], m->x + 2 * (size_t)en[1],
                 q.data() + 4 * (size_t)ec[0], q.data() + 4 * (size_t)ec[1],
                 adt.data() + ec[0], adt.data() + ec[1], f, f + 4);
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int b = 0; b < m->nbedge; ++b) {
      double *f = &bflux[4 * (size_t)b];
      for (int k = 0; k < 4; ++k) f[k] = -0.0;
      const int *bn = m->bedge + 2 * (size_t)b;
      const int cc = m->becell[b];
      k_bres_calc(c, m->x + 2 * (size_t)bn[0], m->x + 2 * (size_t)bn[1],
                  q.data() + 4 * (size_t)cc, adt.data() + cc, f, m->bound + b);
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int n = 0; n < nc; ++n) {
      double *r = &res[4 * (size_t)n];
      for (int j = inc_off[n]; j < inc_off[n + 1]; ++j) {
        const int ent = inc_ent[j];
        const double *f = ent >= 0 ? &eflux[4 * (size_t)ent] : &bflux[4 * (size_t)(-1 - ent)];
        for (int k = 0; k < 4; ++k) r[k] += f[k];
      }
    }
  };

  for (int iter = 1; iter <= niter; ++iter) {
    double t0 = now();
    if (exec == 0) {
      orc_save_soln(nc, q.data(), qold.data());
    } else if (exec == 2) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int n = 0; n < nc; ++n) k_save_soln(&q[4 * (size_t)n], &qold[4 * (size_t)n]);
    } else {
#pragma omp parallel for num_threads(nthreads)
      for (int t = 0; t < nthreads; ++t) {
        int lo = (int)((long long)nc * t / nthreads), hi = (int)((long long)nc * (t + 1) / nthreads);
        for (int n = lo; n < hi; ++n) k_save_soln(&q[4 * (size_t)n], &qold[4 * (size_t)n]);
      }
    }
    tl[0] += now() - t0;
    double rms = 0, maxerr = 0;
    int errloc = 0;
    for (int k = 0; k < 2; ++k) {
      t0 = now();
      if (exec == 0)
        adt_range(m, c, q.data(), adt.data(), 0, nc);
      else if (exec == 2) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int n = 0; n < nc; ++n) adt_range(m, c, q.data(), adt.data(), n, n + 1);
      } else
        run_blocks(adt_plan.p, nthreads,
                   [&](int lo, int hi) { adt_range(m, c, q.data(), adt.data(), lo, hi); });
      double t1 = now();
      tl[1] += t1 - t0;
      if (exec == 0)
        res_range(m, c, q.data(), adt.data(), res.data(), 0, m->nedge);
      else if (exec == 2)
        res_bres_ordered();  // both loops: timed as res_calc
      else
        run_blocks(res_plan.p, nthreads, [&](int lo, int hi) {
          res_range(m, c, q.data(), adt.data(), res.data(), lo, hi);
        });
      double t2 = now();
      tl[2] += t2 - t1;
      if (exec == 0)
        bres_range(m, c, q.data(), adt.data(), res.data(), 0, m->nbedge);
      else if (exec == 1)
        run_blocks(bres_plan.p, nthreads, [&](int lo, int hi) {
          bres_range(m, c, q.data(), adt.data(), res.data(), lo, hi);
        });
      double t3 = now();
      tl[3] += t3 - t2;
      rms = 0.0;
      maxerr = 0.0;
      errloc = 0;
      if (exec == 0) {
        update_range(qold.data(), q.data(), res.data(), adt.data(), &rms,
                     &maxerr, &errloc, 0, nc);
      } else if (exec == 2) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int n = 0; n < nc; ++n) {  // update.h:1-19 per element
          const double adti = 1.0 / adt[n];
          for (int k = 0; k < 4; ++k) {
            const size_t i = 4 * (size_t)n + k;
            const double del = adti * res[i];
            q[i] = qold[i] - del;
            res[i] = 0.0;
            sq[i] = del * del;
          }
        }
        for (int n = 0; n < nc; ++n)
          for (int k = 0; k < 4; ++k) {
            const double sqdel = sq[4 * (size_t)n + k];
            rms += sqdel;
            if (sqdel > maxerr) {
              maxerr = sqdel;
              errloc = n;
            }
          }
      } else {
        // openmp/loop_host.hpp.jinja:183-202,221-245,300-324: per-thread
        // partials 64 elements apart, so no two threads write one cache line
        const int pad = 64;
        std::vector<double> rms_l((size_t)nthreads * pad, 0.0), max_l((size_t)nthreads * pad, maxerr);
        std::vector<int> loc_l((size_t)nthreads * pad, 0);
#pragma omp parallel for num_threads(nthreads)
        for (int t = 0; t < nthreads; ++t) {
          int lo = (int)((long long)nc * t / nthreads), hi = (int)((long long)nc * (t + 1) / nthreads);
          update_range(qold.data(), q.data(), res.data(), adt.data(), &rms_l[(size_t)t * pad],
                       &max_l[(size_t)t * pad], &loc_l[(size_t)t * pad], lo, hi);
        }
        for (int t = 0; t < nthreads; ++t) rms += rms_l[(size_t)t * pad];
        for (int t = 0; t < nthreads; ++t)
          if (max_l[(size_t)t * pad] > maxerr) {
            maxerr = max_l[(size_t)t * pad];
            errloc = loc_l[(size_t)t * pad];
          }
      }
      tl[4] += now() - t3;
      if (k == 0 && rms_k0_raw) rms_k0_raw[iter - 1] = rms;
    }
    if (rms_out) rms_out[iter - 1] = std::sqrt(rms / (double)nc);
    if (maxerr_out) maxerr_out[iter - 1] = maxerr;
    if (errloc_out) errloc_out[iter - 1] = errloc;
  }
  if (q_out) std::memcpy(q_out, q.data(), sizeof(double) * 4 * (size_t)nc);
  if (loop_seconds)
    for (int i = 0; i < 5; ++i) loop_seconds[i] = tl[i];
  return 0;
}
