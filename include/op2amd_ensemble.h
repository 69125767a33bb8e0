/*
 * Batched ensembles of independent airfoil meshes.
 *
 * An ensemble is many airfoil problems of different sizes and free streams
 * advanced in lock step.  All member meshes are concatenated into one set of
 * nodes/cells/edges/bedges; each loop's work is a table of items, runs of one
 * member's elements (tiles of <= 1024 cells, the owner patches of res_calc
 * with bres_calc inside them, or <= 64 boundary cells when bres_calc runs
 * apart), with per-member offsets into it.  Each
 * loop of the time step is ONE persistent launch over every member's items: a
 * grid of as many workgroups as the GPU holds at once, each walking the item
 * table (no edge crosses members, so the concatenation is exact).  update,
 * save_soln and adt_calc of a cell run in one pass (the update of the second
 * half step does the next step's save_soln and adt_calc; one save_soln and
 * adt_calc launch precedes the first step), the per-member rms/maxerr/errloc
 * reductions run through per-tile partials, and the whole time step is
 * captured in one HIP graph.  The baseline launches every loop once per
 * member.  The arithmetic per element is the single-mesh airfoil's
 * (owner-compute res_calc/bres_calc), so each member's trajectory equals its
 * own sequential run (airfoil.cpp:221-297) bit for bit in q.
 *
 * Members: member m (0-based) has scale s = 1 + (max_I/min_I - 1) u, mesh
 * size round(min_I s) x round(min_J s), Mach = mach_lo + (mach_hi - mach_lo) u',
 * alpha = alpha_lo + (alpha_hi - alpha_lo) u'' degrees, with u, u', u'' drawn
 * in that order per member from one SplitMix64 stream seeded with `seed`; its
 * mesh is op2amd_mesh_generate(I, J, seed + 1 + m, ordering).  The free stream
 * is qinf = (r, r u cos a, r u sin a, r e) with the constants of
 * airfoil.cpp:149-159 (gam, gm1, p, r as float literals) and u = sqrt(gam p/r)
 * Mach, e = p/(r gm1) + 0.5 u^2, Mach and a in double.
 */
#ifndef OP2AMD_ENSEMBLE_H
#define OP2AMD_ENSEMBLE_H

#include <stdint.h>

#include "op2amd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nmesh;
  int min_I, min_J, max_I, max_J;
  uint64_t seed;
  double mach_lo, mach_hi, alpha_lo_deg, alpha_hi_deg;
  int ordering; /* OP2AMD_ORDER_* of every member */
} op2amd_ensemble_spec;

typedef struct {
  int I, J, ncell;
  double mach, alpha_deg;
  double qinf[4];
} op2amd_ensemble_member;

/* Fill members[0..nmesh-1].  Returns 0 on success. */
int op2amd_ensemble_describe(const op2amd_ensemble_spec *spec,
                             op2amd_ensemble_member *members);

typedef struct {
  int niter, warmup;
  int batched;     /* 1: one launch per loop for all members; 0: per member */
  int graphed;     /* batched only: time step replayed from a HIP graph */
  int member_timing; /* per-member step times: looped mode, HIP events around
                        each member's launches; batched mode, device-clock
                        stamps of every work item inside the batched step
                        (16 bytes per item per iteration, kept on the device
                        for all niter iterations: about 3 MB per iteration
                        for 256 members of up to 640x320) */
  int autotune;      /* batched only: workgroup sizes of the two update+
                        adt_calc launches and of bres_calc picked by timing
                        the first ten steps (candidates 64..1024, one per
                        step, each twice; a candidate's time is its faster
                        step; the results do not depend on them) */
  int max_workgroups; /* > 0: at most this many workgroups per launch (each
                         then walks more work items); 0: as many as the GPU
                         holds at once */
  int patch_cells;    /* cells per owner patch of res_calc, 64..1024; 0: 256
                         (the results do not depend on it) */
  int separate_bres;  /* 1: bres_calc in launches of its own; 0: inside
                         res_calc's launch where every boundary cell is owned
                         (the same bits) */
} op2amd_ensemble_opts;

typedef struct {
  /* caller-supplied arrays (any may be NULL) */
  double *rms;          /* [niter][nmesh] sqrt(rms/ncell) of the 2nd update */
  double *maxerr;       /* [niter][nmesh] */
  int *errloc;          /* [niter][nmesh] cell index within the member */
  double *q_out;        /* [total cells][4], members in order */
  double *member_step_ms; /* [nmesh] mean step time per member: looped mode
                             with member_timing, the member's launches;
                             batched mode with member_timing, from the step's
                             first work item start to the member's last work
                             item end; batched otherwise: the whole step */
  double *member_busy_ms; /* [nmesh] batched mode with member_timing: summed
                             over the step's launches, the member's last item
                             end minus its first item start */
  /* filled in */
  double seconds;       /* wall time of iterations warmup+1..niter */
  double setup_seconds;
  long long total_cells, total_edges;
  int blocks[4];        /* workgroup sizes: update+adt_calc (first half
                           step), res_calc (its owner patch size), bres_calc
                           (0: fused into res_calc), update+save_soln+adt_calc
                           (second half step) */
  int grids[4];         /* persistent grids (workgroups) of the same launches
                           over all members */
  int dead_res;         /* 1: every cell is owned by a res_calc patch, so the
                           update does not zero res and res_calc does not
                           read it */
  int fused_bres;       /* 1: bres_calc runs inside res_calc's launch, each
                           boundary cell adding its boundary edges after its
                           interior ones (the order of the two loops) */
} op2amd_ensemble_result;

int op2amd_ensemble_run(const op2amd_ensemble_spec *spec,
                        const op2amd_ensemble_opts *opts,
                        op2amd_ensemble_result *res);

#ifdef __cplusplus
}
#endif

#endif
