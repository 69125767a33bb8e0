/*
 * The airfoil application (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:72-315)
 * written against this library's OP2 API and packaged as a callable entry
 * point, so that tests and benchmarks drive the same application code a user
 * would run.  The time step is the reference's: save_soln, then twice
 * adt_calc, res_calc, bres_calc, update with rms/maxerr/errloc reductions
 * (airfoil.cpp:221-297).
 *
 * Differences from the reference app, all optional:
 *   - alpha_mode 1 rotates the free stream by alpha (qinf[1] = r u cos a,
 *     qinf[2] = r u sin a); the reference computes alpha but leaves qinf[2] = 0
 *     (airfoil.cpp:154-164), which on this channel mesh is an exact steady
 *     state;
 *   - the mesh comes from memory instead of ./new_grid.dat;
 *   - execution modes select the backend's plans, synchronisation and HIP-graph
 *     capture of the time step.
 */
#ifndef OP2AMD_AIRFOIL_H
#define OP2AMD_AIRFOIL_H

#include "op2amd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OP2AMD_MODE_PLAIN 0   /* sync + gbl read-back after every loop,
                                 default plans (LDS-staged, coloured), or the
                                 staged/coloured plans the options name (any
                                 other plan: return code 5) */
#define OP2AMD_MODE_TUNED 1   /* per-loop plans from the options, sync only
                                 where a gbl result is needed */
#define OP2AMD_MODE_GRAPHED 2 /* TUNED plans, each time step replayed from a
                                 captured HIP graph; rms history read back once */

typedef struct {
  int niter;        /* total iterations, warm-up included */
  int warmup;       /* leading iterations excluded from `seconds` */
  int alpha_mode;   /* 0 reference free stream, 1 rotated by alpha */
  int mode;         /* OP2AMD_MODE_* */
  int soa;          /* 1: device dats SoA (OP_AUTO_SOA) */
  int plan_res, block_res;     /* OP2H_PLAN_* and block size, 0 = default */
  int plan_bres, block_bres;
  int block_direct;            /* save_soln, update */
  int block_adt;
  int kernel_timing;           /* 1: per-loop HIP event timing */
  int plan_on_device;          /* 1: colour plans on the GPU */
  int renumber;                /* 1: op_renumber(pecell) after declarations,
                                  as airfoil_hdf5/dp/airfoil.cpp:155 does;
                                  results are reported in the input numbering;
                                  2: the same numbering on the device only
                                  (op_init token OP_HIP_RENUMBER=pecell), the
                                  application keeping its own */
  int mixed;                   /* 1: p_adt and p_res declared "float": fluxes
                                  and adt in FP32, q/qold/update/rms in FP64 */
  int autotune;                /* >0: runtime autotuning with this many timed
                                  calls per variant (loops left on AUTO); in
                                  graphed mode iterations run eagerly until
                                  every loop has chosen, then the step is
                                  captured */
  int graph_steps;             /* GRAPHED: time steps per captured graph
                                  (at most 16; 0 or 1: one), lowered to a
                                  divisor of the timed iteration count */
  const char *dump;            /* non-NULL: the final q is written as the
                                  reference app writes it (airfoil.cpp:302-303)
                                  to <dump>.dat (op_print_dat_to_txtfile) and
                                  <dump>.bin (op_print_dat_to_binfile), in the
                                  run's numbering; when renumbered, the cells'
                                  new_of_old goes to <dump>.perm (count, then
                                  one index per line) */
} op2amd_airfoil_opts;

typedef struct {
  /* arrays of niter entries supplied by the caller (any may be NULL) */
  double *rms;       /* sqrt(rms/ncell) after the second update of each iter */
  double *rms_k0;    /* raw rms sum of the first update of each iter */
  double *maxerr;    /* second update */
  int *errloc;
  double *q_out;     /* final q, ncell*4 (AoS) */
  /* filled in */
  double seconds;        /* wall time of iterations warmup+1..niter */
  double setup_seconds;  /* declarations, plans, warm-up */
  double loop_kernel_ms[5];  /* per loop: summed kernel time (kernel_timing) */
  int loop_calls[5];
  double loop_bytes[5];      /* compulsory bytes of one call of each loop */
  int chosen_plan[5], chosen_block[5];  /* autotuner's choices (-1 if none) */
  int eager_iterations;      /* iterations run before graph capture */
  int graph_steps;           /* time steps per graph launch used */
} op2amd_airfoil_result;

/* Run the airfoil time-marching loop on mesh `m`.  Returns 0 on success. */
int op2amd_airfoil_run(const op2amd_mesh *m, const op2amd_airfoil_opts *o,
                       op2amd_airfoil_result *r);

#ifdef __cplusplus
}
#endif

#endif
