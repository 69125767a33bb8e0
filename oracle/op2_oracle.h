/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this library; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py do, and only to check or time against it.
 *
 * A CPU restatement of the reference's airfoil path:
 *   - the five airfoil kernels, operation for operation as the reference
 *     headers write them (apps/c/airfoil/airfoil_plain/dp/{save_soln,adt_calc,
 *     res_calc,bres_calc,update}.h), built without FMA contraction;
 *   - the sequential executor's semantics (op2/include/op_seq.h:27-58,918-1073):
 *     elements in order 0..n-1, OP_INC applied in place, gbl args written in
 *     place;
 *   - the OpenMP executor (translator-v2/resources/templates/cpp/openmp/
 *     loop_host.hpp.jinja:178-358): plan from op_plan_core(OP_STAGE_ALL),
 *     block colours in sequence, the blocks of one colour in parallel,
 *     elements of a block in order; direct loops split the set evenly over
 *     threads; gbl reductions through per-thread partials combined in thread
 *     order;
 *   - op_plan_core(OP_STAGE_ALL) for a single rank
 *     (op2/src/core/op_rt_support.cpp:301-1126): block partition, per-block
 *     unique indirect lists and local maps, greedy first-fit thread and block
 *     colouring with 32-bit masks;
 *   - the airfoil driver (airfoil.cpp:149-297).
 *
 * Parity pins (tests/test_oracle.py): reference-built golden vectors
 * (tests/golden/), bit-for-bit; closed forms and invariants where the golden
 * vectors leave a function open.
 */
#ifndef OP2_ORACLE_H
#define OP2_ORACLE_H

#include "op2amd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double gam, gm1, cfl, eps, mach, alpha, qinf[4];
} orc_consts;

/* Constants of airfoil.cpp:149-164; alpha_mode 1 rotates the free stream by
 * alpha (qinf[1] = r u cos a, qinf[2] = r u sin a), alpha_mode 0 keeps
 * qinf[2] = 0 as the reference does. */
void orc_airfoil_consts(int alpha_mode, orc_consts *c);

/* Single loops in op_seq.h order.  All arrays AoS as the reference host
 * arrays. */
void orc_save_soln(int ncell, const double *q, double *qold);
void orc_adt_calc(const op2amd_mesh *m, const orc_consts *c, const double *q,
                  double *adt);
void orc_res_calc(const op2amd_mesh *m, const orc_consts *c, const double *q,
                  const double *adt, double *res);
void orc_bres_calc(const op2amd_mesh *m, const orc_consts *c, const double *q,
                   const double *adt, double *res);
/* rms/maxerr/errloc are in/out, as gbl args in op_seq.h. */
void orc_update(int ncell, const double *qold, double *q, double *res,
                const double *adt, double *rms, double *maxerr, int *errloc);

/* Full airfoil run (airfoil.cpp:219-297).
 *   exec 0: sequential;  exec 1: OpenMP executor with `nthreads` threads
 *   (<=0: all) and plan part size `part_size` (0: the reference heuristic);
 *   exec 2: the sequential executor's results bit for bit, evaluated with
 *   `nthreads` threads (each edge's increments evaluated once, each cell
 *   adding its own in the sequential order, the update's reductions summed
 *   serially in element order); its res_calc time includes bres_calc.
 * Outputs per iteration (arrays of niter, any may be NULL): the raw rms sum of
 * the first update call, the normalised rms after the second
 * (sqrt(rms/ncell)), maxerr and errloc of the second.  q_out: final q.
 * loop_seconds[5]: wall time spent in save_soln, adt_calc, res_calc,
 * bres_calc, update.  Returns 0 on success. */
int orc_airfoil_run(const op2amd_mesh *m, int niter, int alpha_mode, int exec,
                    int nthreads, int part_size, double *rms_k0_raw,
                    double *rms, double *maxerr, int *errloc, double *q_out,
                    double *loop_seconds);

/* As orc_airfoil_run with the free stream `qinf` given explicitly (the other
 * constants as airfoil.cpp:149-152); the initial q is qinf everywhere. */
int orc_airfoil_run_qinf(const op2amd_mesh *m, int niter, const double *qinf,
                         int exec, int nthreads, int part_size,
                         double *rms_k0_raw, double *rms, double *maxerr,
                         int *errloc, double *q_out, double *loop_seconds);

/* op_plan_core(OP_STAGE_ALL) restatement for one indirect loop on one rank.
 * Argument m (0..nargs-1) uses map table maps[m] (row-major [set_size][mdim[m]],
 * NULL for a direct arg) at index idx[m]; inds[m] >= 0 numbers the distinct
 * indirect dats (-1: direct/gbl); acc[m] is the access code (0 READ, 1 WRITE,
 * 2 RW, 3 INC); dat_bytes[m] the bytes of one dat element; to_size[m] the size
 * of the target set.  part_size 0 selects the reference heuristic.
 */
typedef struct {
  int nblocks, ncolors, bsize;
  int *offset, *nelems, *nthrcol;  /* [nblocks] */
  int *thrcol;                     /* [set_size] */
  int *blk_col;                    /* [nblocks] */
  int *ncolblk;                    /* [ncolors] */
  int *blkmap;                     /* [nblocks] */
  int ninds;
  int *ind_sizes, *ind_offs;       /* [nblocks][ninds] */
  int *ind_map;                    /* [ninds][set_size*k] concatenated */
  int *nindirect;                  /* [ninds] */
  short *loc_map;                  /* [nargs][set_size] (indirect args) */
} orc_plan;

int orc_plan_build(int set_size, int nargs, const int *const *maps,
                   const int *mdim, const int *idx, const int *inds,
                   const int *acc, const int *dat_bytes, const int *to_size,
                   int part_size, orc_plan *out);
void orc_plan_free(orc_plan *p);

#ifdef __cplusplus
}
#endif

#endif
