/*
 * op2_hip_rt.h — the runtime ABI that generated loop stubs call.
 *
 * It replaces what the reference's HIP stubs take from its CUDA runtime
 * (hip/loop_host.hpp.jinja:188-307: op_plan_get_stage, reallocConstArrays,
 * reallocReductArrays, mv{Const,Reduct}ArraysTo{Device,Host}, OP_reduct_h/d,
 * op_mpi_halo_exchanges_grouped, op_mpi_set_dirtybit_cuda, hipDeviceSynchronize):
 *   - loop lifecycle: argument checks, upload of host-dirty dats
 *     (op_cuda_rt_support.cpp:398-452 semantics), dirty marking
 *     (:454-462), per-loop timing and the synchronisation policy;
 *   - device data layout (AoS or SoA with a wave64-padded stride);
 *   - execution plans: the block partition and two-level colouring of
 *     op_plan_core(OP_STAGE_ALL) (op_rt_support.cpp:301-1126), the flat
 *     colouring of OP_COLOR2, and the owner-compute inverse map;
 *   - gbl reduction slots and their delivery to host pointers, immediately or
 *     at the end of a captured HIP graph.
 * All pointers in the plan structs are device pointers unless named *_h.
 */
#ifndef OP2_HIP_RT_H
#define OP2_HIP_RT_H

#include "op2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OP2H_MAX_ARGS 32
#define OP2H_MAX_GBL 8

typedef struct {
  int kernel_idx;
  const char *name;
  op_set set;
  int plan;        /* OP2H_PLAN_* resolved for this call */
  int block_size;  /* threads per workgroup */
  int part_size;   /* elements per plan block (staged plan) */
  int capturing;   /* 1 while a graph is being recorded */
  int tuning;      /* 1 while this call times an autotuning candidate */
  void *stream;    /* hipStream_t */
  double wall0;
  int nargs;
  op_arg *args;
  /* gbl reductions of this call */
  int ngbl;
  int gbl_arg[OP2H_MAX_GBL];
  void *gbl_slot_d[OP2H_MAX_GBL];
  void *gbl_slot_h[OP2H_MAX_GBL];
} op2h_loop_ctx;

/* Loop lifecycle.  op2h_loop_begin checks the arguments (OP_diags > 0),
 * uploads dats whose host copy is newer, resolves plan and block size
 * (op2h_set_loop_plan, else defaults `def_plan`/`def_block`), and starts the
 * timers.  op2h_loop_end marks written dats as device-newer, delivers gbl
 * results (now, or at graph sync when capturing), synchronises per the sync
 * mode and records timing.  `is_indirect` selects the default plan. */
void op2h_loop_begin(op2h_loop_ctx *ctx, int kernel_idx, const char *name,
                     op_set set, int nargs, op_arg *args, int def_plan,
                     int def_block);
void op2h_loop_end(op2h_loop_ctx *ctx);

/* Kernel timing hooks around each launch (no-ops unless enabled). */
void op2h_kernel_timer_start(op2h_loop_ctx *ctx);
void op2h_kernel_timer_stop(op2h_loop_ctx *ctx);

/* Kernel-attached timing events: when kernel timing is on (and the call is
 * neither a tuning trial nor captured), returns 1 with two events in *start
 * and *stop for hipExtLaunchKernel, which stamps them at the first launch's
 * start and the last launch's end, so the time is the kernels' own (no event
 * packets between them); the stub then skips op2h_kernel_timer_start/stop.
 * Otherwise returns 0 with both NULL. */
int op2h_kernel_timer_events(op2h_loop_ctx *ctx, void **start, void **stop);

/* Deferred launches, for loop-chain fusion.  Inside a captured graph a stub
 * may leave its kernel launch pending, op2h_defer(tag, flush, arg), so that
 * the next stub can fold it into its own launch: that stub announces it with
 * op2h_accept_deferred(tag) before op2h_loop_begin and then takes it over
 * with op2h_take_deferred(tag, &arg) (1, with the deferred arg; 0 when
 * nothing is pending, or, after flushing it, when what is pending has
 * another tag).  Everything else flushes the pending launch
 * first (calls flush(arg)): any other op2h_loop_begin, op2h_device_sync,
 * fetches and dumps, graph begin and end, op_renumber, op_exit.  At most one
 * launch is pending.  op2h_fusion_enabled() is 1 when a stub may defer: the
 * call is being captured, fusion is on (op2h_set_fusion, default on; a
 * setting kept across op_init/op_exit) and neither kernel timing, the
 * autotuner, per-loop syncs nor per-loop profiling needs the loop on its
 * own. */
typedef void (*op2h_flush_fn)(void *arg);
void op2h_defer(const char *tag, op2h_flush_fn flush, void *arg);
void op2h_accept_deferred(const char *tag);
int op2h_take_deferred(const char *tag, void **arg);
void op2h_flush_deferred(void);
void op2h_set_fusion(int on);
int op2h_fusion_enabled(void);
/* Streaming or cache-resident: whether the data a loop streams (`bytes`,
 * its element-indexed reads and writes) exceeds the 256 MiB last-level
 * cache, so that the next loop cannot find it there.  op2h_streaming(bytes)
 * is 1 when it does: by size (the default, mode -1), always (1) or never
 * (0), set by op2h_set_streaming or the op_init token
 * OP_HIP_STREAMING=auto|on|off (a setting kept across op_init/op_exit).
 * Stubs choose by it:
 *   - non-temporal hints on a loop's single-use streams (adt_calc's pcell
 *     rows, q and adt): streaming data gains L2 room for what the kernel
 *     does reuse (the gathered nodes), cache-resident data would lose its
 *     place in the last-level cache for the next loop;
 *   - the cells of the update (and of the direct loops fused with it) per
 *     workgroup: cache-resident, contiguous chunks of about three cells per
 *     thread on XCD-ordered workgroups (neighbouring cells and the nodes
 *     they gather in one XCD's L2); streaming, a grid-stride sweep of at
 *     most 2048 workgroups moving through the set together. */
void op2h_set_streaming(int mode);
int op2h_streaming(size_t stream_bytes);
/* Pending launches taken over by a later stub since the library loaded. */
unsigned long long op2h_fused_count(void);

/* A counter advanced whenever the runtime drops its plans (op_exit,
 * op_renumber): data a stub derives from a plan is valid while it holds. */
unsigned op2h_plan_epoch(void);

/* Device layout of a dat: 0 AoS, else the SoA stride in elements. */
int op2h_dat_soa_stride(op_dat dat);

/* Device result slot for gbl arg `argi` of this call (dim*size bytes of the
 * arg, plus the linked info arg for OP_MIN/OP_MAX when `info_argi` >= 0).
 * The kernel writes the loop's own reduction over the set (identity-started)
 * into the slot; op2h_loop_end combines it with the host value. */
void *op2h_gbl_slot(op2h_loop_ctx *ctx, int argi, int info_argi);

/* Deferred gbl reductions inside a captured graph.  A stub that reduces a
 * dim-1 double gbl arg through per-workgroup partials may, while a graph is
 * being captured, hand the partials to the runtime instead of finishing
 * them: op2h_gbl_partials returns 1 with device buffers for `nparts`
 * partials, vals[nparts] and, for OP_MIN/OP_MAX with a linked dim-1 int info
 * arg, locs[nparts] (the info payload, which must be the element index: equal
 * values keep the lowest).  The graph's history kernel then combines them in
 * partial order (thread t takes partials t and t + 1024, then wave trees, then
 * the 16 waves in order) when it records the launch, so one kernel finishes
 * every deferred reduction of a step.  Outside capture it returns 0 with
 * NULLs, and the stub finishes the reduction into op2h_gbl_slot itself. */
int op2h_gbl_partials(op2h_loop_ctx *ctx, int argi, int info_argi, int nparts,
                      double **vals, int **locs);

/* Device scratch for per-workgroup partials and the zero-initialised
 * completion counter of the single-pass (last-workgroup-reduces) reduction.
 * Kernels must leave the counter at zero when they finish. */
void *op2h_reduct_scratch(size_t bytes);
unsigned int *op2h_reduct_counter(void);

/* ---- constants ----
 * op_decl_const_char offers each constant to every registered hook; a hook
 * copies the constants it knows (by name) to its device copies and returns 1,
 * or returns 0.  Generated kernel files register one hook each at
 * load time.  A name no hook knows is fatal, as in the reference
 * (hip/master_kernel.cpp.jinja:21-25). */
typedef int (*op2h_const_hook)(int dim, const char *type, int size,
                               const char *dat, const char *name);
void op2h_register_const_hook(op2h_const_hook hook);

/* Begin a loop whose stub implements only the plans in `plan_mask` (bit
 * 1 << OP2H_PLAN_*): the autotuner tries those; a request for another plan
 * made for this loop by name is fatal, one made for all loops is ignored.  op2h_loop_begin(ctx, ...) is op2h_loop_begin_mask
 * with every plan when def_plan is OP2H_PLAN_STAGED, else only def_plan. */
void op2h_loop_begin_mask(op2h_loop_ctx *ctx, int kernel_idx, const char *name,
                          op_set set, int nargs, op_arg *args, int def_plan,
                          int def_block, unsigned plan_mask);

/* ---- execution plans ---- */

/* Block partition + two-level colouring (OP_STAGE_ALL semantics), with
 * indirect staging lists deduplicated per (map) rather than per dat. */
typedef struct {
  int set_size, part_size, nblocks, ncolors;
  int nmaps;                   /* distinct maps among indirect args */
  int *ncolblk_h;              /* [ncolors] blocks per colour (host) */
  int *colstart_h;             /* [ncolors+1] first index into blkmap (host) */
  int *max_nthrcol_h;          /* [1] */
  int max_ind_h[OP2H_MAX_ARGS];/* per map: max unique targets in a block */
  int *blkmap;                 /* [nblocks] blocks ordered by colour */
  int *offset, *nelems;        /* [nblocks] */
  int *nthrcol;                /* [nblocks] */
  unsigned char *thrcol;       /* [set_size] thread colour within its block */
  int *ind_offs[OP2H_MAX_ARGS];  /* per map: [nblocks] start in ind_map */
  int *ind_sizes[OP2H_MAX_ARGS]; /* per map: [nblocks] unique targets */
  int *ind_map[OP2H_MAX_ARGS];   /* per map: concatenated unique targets */
  unsigned short *loc_map[OP2H_MAX_ARGS]; /* per (map, map index) in arg
                                   order of first use: [set_size] local ids */
  int map_of_arg[OP2H_MAX_ARGS]; /* arg -> staged map slot, -1 if direct */
  int loc_of_arg[OP2H_MAX_ARGS]; /* arg -> loc_map slot, -1 if direct */
  int built_on_device;         /* 1 if the colouring was computed on the GPU */
  double build_ms;
} op2h_staged_plan;

/* Flat colouring of all elements (OP_COLOR2): elements permuted by colour. */
typedef struct {
  int set_size, ncolors;
  int *col_offsets_h;  /* [ncolors+1] */
  int *col_reord;      /* [set_size] element ids grouped by colour */
  int built_on_device;
  double build_ms;
} op2h_color_plan;

/* Owner-compute inverse of an increment map: for each target element, the
 * (element * dim + slot) entries that increment it, in ascending element
 * order.  `targets` lists the target elements with at least one entry. */
typedef struct {
  int from_size, to_size, dim, ntargets;
  int max_count;  /* most entries of one target */
  int *offsets;   /* [ntargets+1] */
  int *entries;   /* [from_size*dim] element*dim+slot */
  int *targets;   /* [ntargets] */
  double build_ms;
} op2h_gather_plan;

/* Owner-compute partition of a loop whose increments all go through one map
 * into one dat (the "owner" set is that map's target set).  Targets are cut
 * into blocks of at most `block_cells` by breadth-first growth from the
 * lowest-numbered unassigned target (so blocks are compact patches whatever
 * the numbering); every element with an increment target in a block is an
 * "instance" of that block (elements on a block boundary belong to two).
 * Per block: the instances in ascending element order, the owned targets in
 * ascending order, and per owned target its increments as ushort entries
 * `local instance * ninc + k` (k = position of the increment among the loop's
 * increment args) in ascending order, i.e. the order of the sequential
 * executor.  A stub evaluates each instance once (reading its indirect data
 * through the element's map rows) and sums each owned target's increments in
 * list order. */
typedef struct {
  int set_size, to_size, block_cells, nblocks;
  int ninc;                    /* increment args through the owner map */
  int total;                   /* instances over all blocks */
  int max_elems, max_owned;    /* per-block maxima */
  int *owned_offs;             /* [nblocks+1] */
  int *owned;                  /* owned targets, block by block */
  int *elem_offs;              /* [nblocks+1] */
  int *elems;                  /* [total] element of each instance */
  int *csr_offs;               /* [n_owned + 1], in `owned` order */
  unsigned short *csr;         /* increment entries */
  int inc_of_arg[OP2H_MAX_ARGS]; /* arg -> k among increments, -1 if none */
  double redundancy;           /* total / set_size */
  double build_ms;
  int n_owned;                 /* owned targets over all blocks */
  int max_incs;                /* most increments of one owned target */
  const unsigned short *csr4;  /* [n_owned][4]: each owned target's
                                  increments as a fixed-width list (0xFFFF
                                  pads), when max_incs <= 4; else NULL */
} op2h_owner_plan;

const op2h_staged_plan *op2h_plan_staged(op2h_loop_ctx *ctx);
const op2h_color_plan *op2h_plan_color(op2h_loop_ctx *ctx);
const op2h_gather_plan *op2h_plan_gather(op_map map);
/* Owner plan with blocks of ctx->part_size targets (NULL and a message if the
 * loop's increments do not all go through one map into one dat, or a block
 * outgrows the 16-bit local ids). */
const op2h_owner_plan *op2h_plan_owner(op2h_loop_ctx *ctx);
const op2h_owner_plan *op2h_plan_owner_for(const char *name, op_set set,
                                           int block_cells, int nargs,
                                           op_arg *args);
/* As op2h_plan_owner, but under transparent renumbering a block's instances
 * stay in the device's element order (each target's increments still follow
 * the application's): for stubs that evaluate all of a block's instances
 * before summing any increment, where the device order keeps neighbouring
 * elements in neighbouring lanes.  Without transparent renumbering it is
 * op2h_plan_owner. */
const op2h_owner_plan *op2h_plan_owner_local(op2h_loop_ctx *ctx);
/* Copy an owner plan's block structure back to the host (any pointer may be
 * NULL): owned_offs [nblocks+1], owned, elem_offs [nblocks+1], elems,
 * csr_offs, csr.  Returns 0. */
int op2h_owner_plan_download(const op2h_owner_plan *p, int *owned_offs,
                             int *owned, int *elem_offs, int *elems,
                             int *csr_offs, unsigned short *csr);
/* The partition alone, on the host: targets [0,nto) of a map (nel x dim,
 * row-major) joined by the map indices `inc_idx[0..ninc)` of each element;
 * block_of[t] gets the block of each target (-1 for targets no element
 * references).  Returns the number of blocks, or -1 on bad arguments. */
int op2h_owner_partition(int nto, int nel, int dim, const int *map, int ninc,
                         const int *inc_idx, int block_cells, int *block_of);

/* Reverse Cuthill-McKee order of an undirected graph in CSR form (off[n+1],
 * adj[off[n]], each neighbour list strictly ascending, no entry outside
 * [0,n)): new_of_old[v] is the new number of vertex v.  The rule is the one
 * op_renumber applies to its base set (op_renumber.cpp:224-505 orders that set
 * with Scotch's Gibbs-Poole-Stockmeyer instead): roots in (degree, index)
 * order, each improved by up to 16 George-Liu pseudo-peripheral sweeps from
 * the lowest-(degree, index) vertex of the last level; neighbours queued by
 * (degree, index); components one after another; the order reversed.
 * on_gpu = 0 builds it on the host, 1 on the current device, level by level,
 * and 2 on the device with the two-pass level sort that is otherwise used only
 * when a packed 64-bit (parent, degree, index) key is too narrow; all give the
 * same permutation.  Returns 0, or -1 on bad arguments. */
int op2h_rcm_order(int n, const int *off, const int *adj, int on_gpu,
                   int *new_of_old);

/* Copy a plan's colouring back to the host for checking: thread colours,
 * block colours (per block), and block offsets/sizes.  Returns 0. */
int op2h_staged_plan_download(const op2h_staged_plan *p, int *thrcol,
                              int *blkcol, int *offset, int *nelems);
int op2h_color_plan_download(const op2h_color_plan *p, int *color_of_elem);

/* Build the plans of a loop without running it (for inspection and tests):
 * descriptors as in op_par_loop, on loop name `name`. */
const op2h_staged_plan *op2h_plan_staged_for(const char *name, op_set set,
                                             int part_size, int nargs,
                                             op_arg *args);
const op2h_color_plan *op2h_plan_color_for(const char *name, op_set set,
                                           int nargs, op_arg *args);

/* Device buffer of `bytes` private to a generated stub (e.g. a loop-specific
 * packing of a plan), cached under `key` with the plans: freed by op_exit,
 * dropped by op_renumber.  *fresh is 1 when the caller must fill it (before
 * any graph capture). */
void *op2h_plan_user(const char *key, size_t bytes, int *fresh);

/* Under transparent renumbering (op_init token OP_HIP_RENUMBER or
 * OP_HIP_RENUMBER=<map>): before the first loop the runtime renumbers the
 * sets as op_renumber(<map>) would (default map: the first declared one of
 * dimension >= 2 from the largest set), while the application keeps its own
 * numbering: op_fetch_data, op_fetch_data_idx and the op_print_dat_to_*
 * dumps give rows in the application's order, op_arg_idx and op_arg_info
 * payloads carry the application's indices, and increments that the
 * sequential order fixes (owner and gather plans) keep the application's
 * element order, so results are those of the unrenumbered run.  An explicit
 * op_renumber call is then a no-op.  This returns, on the device, the
 * application's index of each element of `set` (old_of_new), or NULL when
 * the set is not renumbered that way. */
const int *op2h_set_old_index(op_set set);

/* Compulsory ("algorithmic") bytes of one call of a loop: every dat element
 * and map entry the loop touches, counted once; RW and INC count a read and a
 * write.  Indirectly accessed dats count the distinct targets of their map. */
double op2h_algorithmic_bytes(op_set set, int nargs, const op_arg *args);

/* Where plans are coloured: 1 on the GPU (default), 0 on the host. */
void op2h_set_plan_builder(int on_device);

/* ---- streaming ceiling (measurement) ----
 * The HBM streaming rate of copy kernels of several shapes, the ceiling the
 * loops' achieved bandwidth is read against beside the 8 TB/s peak.
 * variant: 0 one 16-byte vector per thread; 1 four per thread; 2 four per
 * thread, non-temporal loads and stores; 3 eight per thread; 4 a persistent
 * grid (8 workgroups per CU) striding four vectors per iteration; 5 read
 * only (four per thread); 6 write only (four per thread).  bytes: size of the
 * source and of the destination buffer (at least 1 MiB; allocated here).
 * gbs: bytes moved (read + written) per second / 1e9 over the fastest of
 * `reps` timed calls; us_per_call (may be NULL): that call's time.
 * Returns 0, -1 on bad arguments, -2 on a HIP error. */
#define OP2H_STREAM_VARIANTS 7
int op2h_stream_probe(size_t bytes, int variant, int reps, double *gbs,
                      double *us_per_call);

/* ---- the update loop without allocating ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the update loop of op_par_loop_airfoil_5_update (update.h:1-19)
 * on caller-owned device arrays of ncell cells (qold, q, res: 4 doubles per
 * cell, AoS when soa_stride is 0, else SoA with that stride; adt: 1 double),
 * with the same kernels, so the same bits for the same block size.  The
 * loop's own reductions over the cells (identity-started: rms sum 0, maxerr
 * -infinity) go to the caller's device struct `result`, on `stream` (NULL:
 * the default stream); the caller combines them with its own values.  It
 * allocates nothing: its per-workgroup partials (at most 2048, 20 bytes each)
 * live in a 40 KiB device array of the library, one per device, shared by all
 * calls on that device — calls that overlap on one device race.  No memset
 * and no host synchronisation, so the call can be captured into a graph.
 * block: 64..1024, a multiple of 64; ncell: any non-negative int.
 * Returns 0, or a negative code after printing the reason: -1 bad arguments,
 * -2 no current device or its scratch address, -3 partials beyond the
 * scratch, -4 launch failure. */
typedef struct {
  double rms;    /* sum over the cells of del^2 */
  double maxerr; /* the largest del^2 */
  int errloc;    /* the lowest cell reaching it */
  int pad;
} op2h_update_result;
int op2h_update_noalloc(int ncell, int soa_stride, int block, const double *qold,
                        double *q, double *res, const double *adt,
                        op2h_update_result *result, void *stream);

/* ---- plan checks (op_plan_check, op2/src/core/op_rt_support.cpp:129-295) ----
 * With OP_diags > 0 the runtime checks every plan it builds on the host and
 * stops the program (exit(-1), after the messages) when a check fails: a
 * plan with a colour conflict or a missing increment gives wrong sums
 * without any other sign.  The checks are the reference's (element totals,
 * block offsets, block permutation, staging-list offsets and order, local ids
 * leading back to the map) plus the colouring checks its stub leaves empty
 * (op_rt_support.cpp:290-294), and the analogous ones for the owner-compute
 * and gather plans.  Each check below works on a host image of a plan,
 * returns the number of failed checks and prints
 * " *** OP_plan_check: <what> error" for each (OP_diags > 6: "<what> OK" for
 * those that pass, as the reference does). */

/* One indirect access of a loop, for the conflict checks: element e touches
 * target table[e*dim + idx] (in [0,to_size)) of the loop's dat number `dat`
 * (dats numbered from 0 within the loop); writes = 1 for OP_INC and OP_RW. */
typedef struct {
  const int *table;
  int dim, idx, to_size, dat, writes;
} op2h_check_access;

/* Host image of a staged (two-level coloured) plan: the fields of
 * op2h_staged_plan, with colstart[ncolors+1] the first blkmap position of
 * each block colour and, per map slot, the map's table, dim and target set
 * size; per loc slot its map slot and map index. */
typedef struct {
  int set_size, nblocks, ncolors;
  const int *colstart, *blkmap, *offset, *nelems, *nthrcol;
  const unsigned char *thrcol;
  int nmaps;
  const int *map_table[OP2H_MAX_ARGS];
  int map_dim[OP2H_MAX_ARGS], map_to_size[OP2H_MAX_ARGS];
  const int *ind_offs[OP2H_MAX_ARGS], *ind_sizes[OP2H_MAX_ARGS], *ind_map[OP2H_MAX_ARGS];
  int nlocs;
  int loc_map_slot[OP2H_MAX_ARGS], loc_idx[OP2H_MAX_ARGS];
  const unsigned short *loc_map[OP2H_MAX_ARGS];
} op2h_staged_plan_host;
int op2h_check_staged_plan(const op2h_staged_plan_host *p, int nacc,
                           const op2h_check_access *acc);

/* Flat colouring: col_offsets[ncolors+1], col_reord[set_size]. */
int op2h_check_color_plan(int set_size, int ncolors, const int *col_offsets,
                          const int *col_reord, int nacc,
                          const op2h_check_access *acc);

/* Host image of an owner-compute plan (op2h_owner_plan) with the increment
 * map's table (set_size x map_dim) and the map index of each increment;
 * csr4 may be NULL.  elem_key (may be NULL): the element order each
 * target's increments follow is ascending elem_key[element] (the
 * application's numbering
 * under transparent renumbering) instead of ascending element. */
typedef struct {
  int set_size, to_size, nblocks, ninc;
  const int *map_table;
  int map_dim;
  int inc_idx[OP2H_MAX_ARGS];
  const int *owned_offs, *owned, *elem_offs, *elems, *csr_offs;
  const unsigned short *csr, *csr4;
  const int *elem_key;
  int instances_by_key; /* 1: a block's instances follow elem_key's order
                           too (else ascending element) */
} op2h_owner_plan_host;
int op2h_check_owner_plan(const op2h_owner_plan_host *p);

/* Gather plan (op2h_gather_plan) of a map (from_size x dim, row-major);
 * elem_key as for the owner plan. */
int op2h_check_gather_plan(int from_size, int to_size, int dim, const int *map,
                           int ntargets, const int *offsets, const int *entries,
                           const int *targets, const int *elem_key);

#ifdef __cplusplus
}
#endif

#endif
