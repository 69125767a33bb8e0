/*
 * Synthetic op_par_loops for plan-selection studies (cost model training).
 *
 * A descriptor names a loop over one of the generator's meshes
 * (op2amd_mesh.h): arity 2 loops over edges through edge->cell; arity 1, 3
 * and 4 loop over cells through the first `arity` columns of cell->node.  The
 * loop reads a direct dat V (dim D) on its set and an indirect dat R (dim D)
 * through every map slot, runs a chain of k multiply-adds, and then
 *   READ : writes the result to a direct dat;
 *   INC  : adds it to an indirect dat W through every map slot (OP_INC);
 *   RW   : W = 0.5 W + v through every map slot (OP_RW).
 * Sets, maps and dats are declared through the OP2 C API and the plans are
 * the runtime's own (op2_hip_rt.h): direct, atomic, flat colour, LDS-staged
 * two-level colour, owner-compute gather.  Each variant is timed with HIP
 * events on the library's stream.
 */
#ifndef OP2AMD_SYNTH_H
#define OP2AMD_SYNTH_H

#include <stdint.h>

#include "op2amd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OP2AMD_SYNTH_DIRECT 0 /* plan id for READ loops (no conflicts) */

typedef struct {
  int I, J;       /* generator mesh size */
  int ordering;   /* OP2AMD_ORDER_* */
  int arity;      /* 1..4 */
  int dim;        /* 1, 2, 4 or 8 */
  int access;     /* OP_READ, OP_INC or OP_RW */
  int is_double;  /* 1: double dats and arithmetic, 0: float */
  int k_fma;      /* multiply-adds per element */
  uint64_t seed;  /* mesh jitter/permutation seed */
} op2amd_synth_desc;

/* Loop shape facts used as cost-model features. */
typedef struct {
  int set_size, target_size;
  double mean_bandwidth;  /* mean |t_a - t_b| over an element's targets */
} op2amd_synth_shape;

/* Time `nvar` variants (plan[i] in OP2H_PLAN_* or OP2AMD_SYNTH_DIRECT, block[i]
 * threads) of the loop: one untimed call (plan build), then `reps` samples of
 * 8 back-to-back calls each; ms[i] is the fastest sample's mean per call.  A variant that does not apply to the access
 * mode (atomic for RW, anything but direct for READ) gets ms[i] = -1.
 * Returns 0 on success. */
int op2amd_synth_time(const op2amd_synth_desc *d, int nvar, const int *plan,
                      const int *block, int reps, double *ms,
                      op2amd_synth_shape *shape);

/* Run variant (plan, block) once from fixed seeded dat values and return the
 * final written dat (W for INC/RW, the direct output for READ), as doubles,
 * in `out` (target_size*dim or set_size*dim values).  For tests. */
int op2amd_synth_run_once(const op2amd_synth_desc *d, int plan, int block,
                          double *out, int max_out);

/* Agreement of plans: from the same seeded dat values, run the reference
 * variant (ref_plan, ref_block) once, then each of the `nvar` variants once,
 * and compare each result (W for INC/RW, the direct output for READ) with the
 * reference's on the device: max_rel_dev[i] = max |w - w_ref| / max |w_ref|
 * (NaN if any difference is NaN), mismatches[i] = elements whose bits differ;
 * both -1 for a variant that does not apply.  `mesh` (may be NULL: generated
 * from the descriptor) is a mesh of the descriptor's shape, read only.
 * Returns 0, 1 for a bad descriptor, 2 if the reference variant does not
 * apply. */
int op2amd_synth_check(const op2amd_synth_desc *d, const op2amd_mesh *mesh, int ref_plan,
                       int ref_block, int nvar, const int *plan, const int *block,
                       double *max_rel_dev, long long *mismatches);

#ifdef __cplusplus
}
#endif

#endif
