/*
 * op2_hip.h — public C ABI of the MI355X-native OP2 backend (libop2hip.so).
 *
 * A translated C++ OP2 application includes this header instead of the
 * reference's op_lib_cpp.h and links against libop2hip.so.  The types, access
 * codes and API functions keep the reference's names, argument meaning and
 * struct layouts, so application code and generated loop stubs written for the
 * reference compile and link unchanged:
 *   - codes and structs:     op2/include/op_lib_core.h:108-211
 *   - C API:                 op2/include/op_lib_c.h:75-121
 *   - C++ templates:         op2/include/op_lib_cpp.h:127-203
 *   - profiling C API:       op2/include/op_profile.h:197-212
 *   - single-rank MPI no-ops op2/src/core/op_dummy_singlenode.cpp:43-208
 *   - generated loop stubs:  translator-v2/resources/templates/cpp/
 *                            loop_host.hpp.jinja:34-41 (one per op_par_loop)
 *
 * Errors follow the reference: fatal conditions print a message and exit(-1);
 * argument checks (OP_diags > 0) print and continue (op_lib_core.cpp:748-829).
 * One host thread drives the library; all device work is issued on one HIP
 * stream owned by the library.
 */
#ifndef OP2_HIP_H
#define OP2_HIP_H

/* the C headers the reference's op_lib_core.h brings into every OP2
 * program (op_lib_core.h:43-50), so that programs relying on them build */
#include <math.h>
#include <stdarg.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

/* access codes, argument types, staging modes: op_lib_core.h:108-125 */
#define OP_READ 0
#define OP_WRITE 1
#define OP_RW 2
#define OP_INC 3
#define OP_MIN 4
#define OP_MAX 5

/* Helper macros user kernels use (op_lib_core.h:245-250). */
#ifndef MIN
#define MIN(a, b) ((a < b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a > b) ? (a) : (b))
#endif
#define OP_WORK 6

#define OP_ARG_GBL 0
#define OP_ARG_DAT 1
#define OP_ARG_IDX 2
#define OP_ARG_INFO 3

#define OP_STAGE_NONE 0
#define OP_STAGE_INC 1
#define OP_STAGE_ALL 2
#define OP_STAGE_PERMUTE 3
#define OP_COLOR2 4

typedef int op_access;
typedef int op_arg_type;
typedef int idx_l_t;
typedef long long idx_g_t;

/* op_lib_core.h:134-147 */
typedef struct {
  int index;
  idx_l_t size;
  char const *name;
  int core_size;
  int exec_size;
  int nonexec_size;
  int stride; /* device SoA stride of dats on this set (elements) */
} op_set_core;
typedef op_set_core *op_set;

/* op_lib_core.h:149-161 */
typedef struct {
  int index;
  op_set from, to;
  int dim;
  int *map;         /* host table, row-major [from->size][dim] */
  idx_g_t *map_gbl; /* unused on one rank */
  int *map_d;       /* device table, same row-major layout */
  char const *name;
  int user_managed;
  bool force_part;
} op_map_core;
typedef op_map_core *op_map;

/* op_lib_core.h:163-179 */
typedef struct {
  int index;
  op_set set;
  int dim, size;    /* size: bytes of one element (dim * sizeof(type)) */
  char *data;       /* host copy, AoS */
  char *data_d;     /* device copy: AoS, or SoA with stride set->stride */
  char const *type; /* "double", "double:soa", "float", "int", ... */
  char const *name;
  char *buffer_d;
  char *buffer_d_r;
  int dirtybit;
  int dirty_hd; /* 0 in sync, 1 host newer, 2 device newer */
  int user_managed;
  void *mpi_buffer;
} op_dat_core;
typedef op_dat_core *op_dat;

/* op_lib_core.h:181-199 */
typedef struct {
  int index;
  op_dat dat;
  op_map map;
  int dim, idx, size;
  char *data, *data_d;
  int *map_data, *map_data_d;
  char const *type;
  op_access acc;
  op_arg_type argtype;
  int sent;
  int opt;
} op_arg;

/* op_lib_core.h:201-211 */
typedef struct {
  char const *name;
  int count;
  double time;
  double *times;
  int ntimes;
  float plan_time;
  float transfer;
  float transfer2;
  double mpi_time;
} op_kernel;

#define OP_ID (op_map) NULL
#define OP_GBL (op_map) NULL

#ifdef __cplusplus
extern "C" {
#endif

/* ---- global state (op_lib_c.h:54-64) ---- */
extern int OP_diags, OP_part_size, OP_block_size, OP_auto_soa, OP_hybrid_gpu;
extern int OP_cache_line_size, OP_maps_base_index;
extern int OP_kern_max, OP_kern_curr;
extern op_kernel *OP_kernels;

/* ---- lifecycle and declarations (op_lib_c.h:75-92) ---- */
/* op_init parses the reference's argv tokens (op_lib_core.cpp:205-281):
 * OP_BLOCK_SIZE=, OP_PART_SIZE=, OP_CACHE_LINE_SIZE=, OP_AUTO_SOA,
 * OP_MAPS_BASE_INDEX=, and this backend's own OP_HIP_PLAN=<loop>:<plan>[:<bs>]
 * (plan one of atomic, color, staged, gather, auto). */
void op_init(int argc, char **argv, int diags);
void op_init_soa(int argc, char **argv, int diags, int soa);
void op_exit(void);
op_set op_decl_set(idx_l_t size, char const *name);
op_map op_decl_map(op_set from, op_set to, int dim, idx_l_t *imap,
                   char const *name);
op_map op_decl_map_long(op_set from, op_set to, int dim, idx_g_t *imap,
                        char const *name);
op_dat op_decl_dat_char(op_set set, int dim, char const *type, int size,
                        char *data, char const *name);
op_dat op_decl_dat_temp_char(op_set set, int dim, char const *type, int size,
                             char const *name);
int op_free_dat_temp_char(op_dat dat);
/* Generated per application (one symbol per op_decl_const in the app): copies
 * the host constant to its device __constant__ copy. */
void op_decl_const_char(int dim, char const *type, int size, char *dat,
                        char const *name);

/* ---- argument constructors (op_lib_c.h:94-102, op_lib_core.cpp:832-1060) ---- */
op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, char const *type,
                  op_access acc);
op_arg op_opt_arg_dat(int opt, op_dat dat, int idx, op_map map, int dim,
                      char const *type, op_access acc);
op_arg op_arg_gbl_char(char *data, int dim, const char *type, int size,
                       op_access acc);
op_arg op_opt_arg_gbl_char(int opt, char *data, int dim, const char *type,
                           int size, op_access acc);
op_arg op_arg_info_char(char *data, int dim, const char *type, int size,
                        int ref);
op_arg op_arg_idx(int idx, op_map map);
void op_arg_check(op_set set, int m, op_arg arg, int *ninds, const char *name);

/* ---- data access and output (op_lib_c.h:104-121, op_lib_core.cpp:1283-1369) ---- */
void op_fetch_data_char(op_dat dat, char *usr_ptr);
void op_fetch_data_idx_char(op_dat dat, char *usr_ptr, int low, int high);
void op_upload_dat(op_dat dat);
void op_download_dat(op_dat dat);
void op_upload_all(void);
idx_g_t op_get_size(op_set set);
void op_print_dat_to_binfile(op_dat dat, const char *file_name);
void op_print_dat_to_txtfile(op_dat dat, const char *file_name);

/* ---- timing, diagnostics (op_lib_core.cpp:1067-1167) ---- */
void op_timers(double *cpu, double *et);
void op_timers_core(double *cpu, double *et);
void op_timing_realloc(int kernel);
void op_timing_output(void);
void op_printf(const char *format, ...);
void op_diagnostic_output(void);
int op_is_root(void);
void op_rank(int *rank);

/* ---- profiling C API (op_profile.h:197-212): a timing tree with levels and
 * JSON output, documented in op_profile.h ---- */
void op_profile_start(const char *name);
void op_profile_enter(const char *name);
void op_profile_enter_kernel(const char *name, const char *target,
                             const char *variant);
void op_profile_next(const char *name);
void op_profile_exit(void);
void op_profile_end(void);
void op_profile_output(void);
void op_profile_output_json(const char *filename);

/* ---- host memory (op_lib_core.h:423-426, op_lib_core.cpp:1388-1420) ----
 * op_malloc returns memory aligned to OP2_ALIGNMENT bytes (its size rounded up
 * to a multiple of it; size 0 gives malloc(0)); op_calloc zeroes it;
 * op_realloc keeps the alignment; all of it may be released with op_free or
 * free. */
#define OP2_ALIGNMENT 64
void *op_malloc(size_t size);
void *op_calloc(size_t num, size_t size);
void *op_realloc(void *ptr, size_t size);
void op_free(void *ptr);

/* ---- renumbering (op2/src/externlib/op_renumber.cpp:224-505) ----
 * Reorders every set reachable from base->to for locality: base->to by
 * reverse Cuthill-McKee on the graph "two elements share a base->from
 * element" (the reference uses Scotch's ordering, not available here), the
 * other sets by the reference's propagation rules (op_renumber.cpp:72-151).
 * Host and device data, maps and plans are all moved to the new numbering,
 * which the application then sees (as after the reference's op_renumber). */
void op_renumber(op_map base);
/* The composed new_of_old permutation applied to `set` by op_renumber calls
 * (identity if none); returns 1 if the set was renumbered. */
int op2h_set_permutation(op_set set, int *new_of_old);

/* ---- single-rank no-ops kept for link compatibility (op_dummy_singlenode.cpp:43-208) ---- */
void op_partition(const char *lib_name, const char *lib_routine,
                  op_set prime_set, op_map prime_map, op_dat coords);
int op_mpi_halo_exchanges(op_set set, int nargs, op_arg *args);
int op_mpi_halo_exchanges_cuda(op_set set, int nargs, op_arg *args);
int op_mpi_halo_exchanges_grouped(op_set set, int nargs, op_arg *args,
                                  int device);
void op_mpi_set_dirtybit(int nargs, op_arg *args);
void op_mpi_set_dirtybit_cuda(int nargs, op_arg *args);
void op_mpi_wait_all(int nargs, op_arg *args);
void op_mpi_wait_all_cuda(int nargs, op_arg *args);
void op_mpi_wait_all_grouped(int nargs, op_arg *args, int device);
void op_mpi_reduce_float(op_arg *args, float *data);
void op_mpi_reduce_double(op_arg *args, double *data);
void op_mpi_reduce_int(op_arg *args, int *data);
void op_mpi_reduce_bool(op_arg *args, bool *data);
void op_mpi_barrier(void);

/* =========================================================================
 * Backend controls (no counterpart in the reference; they replace its
 * compile-time OP_BLOCK_SIZE_<k>/OP_PART_SIZE_<k> macros and translator JSON
 * overrides, hip/loop_host.hpp.jinja:201-206,253-257, scheme.py:38-46).
 * ========================================================================= */

/* Plans for a loop with indirect increments. */
#define OP2H_PLAN_AUTO 0   /* tuned choice if recorded, else OP2H_PLAN_STAGED */
#define OP2H_PLAN_ATOMIC 1 /* one thread per element, FP64 atomics */
#define OP2H_PLAN_COLOR 2  /* global element colouring, one launch per colour */
#define OP2H_PLAN_STAGED 3 /* block partition, LDS staging, two-level colouring */
#define OP2H_PLAN_GATHER 4 /* owner-compute: one thread per target element,
                              increments summed in element order (deterministic,
                              same order as the sequential executor) */
#define OP2H_PLAN_OWNER 5  /* owner-compute partition: the target set is cut
                              into compact blocks; a workgroup stages its
                              block's indirect data in LDS, evaluates each
                              element touching the block once, then sums every
                              owned target's increments in element order
                              (deterministic, sequential-executor order) */
#define OP2H_PLAN_COUNT 6

/* Synchronisation modes. */
#define OP2H_SYNC_EVERY_LOOP 0 /* reference behaviour: device sync and gbl
                                  read-back at the end of every op_par_loop */
#define OP2H_SYNC_GBL_ONLY 1   /* sync only when a loop returns a gbl result */

/* Set the plan and block size used for loop `name` (NULL: all loops).
 * block_size 0 keeps the default.  Returns 0, or -1 for an unknown plan. */
int op2h_set_loop_plan(const char *name, int plan, int block_size);
int op2h_get_loop_plan(const char *name, int *plan, int *block_size);
void op2h_set_sync_mode(int mode);
int op2h_get_sync_mode(void);

/* Online autotuning.  With autotuning on, every loop whose plan is AUTO runs
 * its first calls through each candidate variant — (plan, block size) for
 * loops with indirect increments, block size for the others — `trials` calls
 * each, timed with HIP events, then keeps the fastest.  Tuning happens in
 * eager calls only; a loop still tuning when a graph is captured is recorded
 * with its best variant so far.  Each timed call is kept as a record. */
typedef struct {
  char loop[32];
  int set_size;
  int plan, block_size;
  double kernel_ms;   /* one call */
} op2h_tuning_record;
void op2h_set_autotune(int on, int trials);
int op2h_tuning_done(void);  /* 1 when every seen loop has chosen */
int op2h_tuning_records(int max, op2h_tuning_record *out);
int op2h_tuning_choice(const char *name, int *plan, int *block_size);

/* Plan choice by a cost model (the library's MLPs, op2amd_costmodel.h) for
 * loops left on AUTO: at a loop's first call the runtime describes the loop
 * (set and target sizes, map arity and locality, dat dim and precision,
 * access, arithmetic per element), builds the feature vector of every
 * (plan, block) variant the loop implements, evaluates the members on the
 * device and ranks the variants by their mean prediction.  top = 1 takes the
 * best; top > 1 times the `top` best on the loop's next eager calls (the
 * autotuner's trials) and keeps the fastest.  Features are standardised as
 * (x - mean) / scale, nfeat values each (OP2H_MODEL_FEATURES).  n_members = 0
 * removes the model.  The model stays installed across op_init/op_exit.
 * Returns 0, or -1 on bad arguments. */
#define OP2H_MODEL_FEATURES 22
struct op2amd_mlp;
int op2h_set_plan_model(struct op2amd_mlp *const *members, int n_members, int nfeat,
                        const float *mean, const float *scale, int top);
/* Arithmetic per element of loop `name`'s kernel, in multiply-add
 * equivalents: a cost-model feature (16 when not given; clamped to 4-64,
 * the range the model is trained on). */
void op2h_set_loop_arith(const char *name, int per_element);
/* The description the cost model used for loop `name` (set by its first
 * call with a model installed): set size, target size, arity, dim,
 * is_double, arithmetic, mean map bandwidth, access (0 read, 1 inc, 2 rw).
 * Returns 0, or -1 when there is none. */
int op2h_plan_model_loop(const char *name, double *desc8);

/* Step capture into a HIP graph.  Between op2h_graph_begin() and
 * op2h_graph_end() every op_par_loop is recorded instead of run; gbl
 * reduction results of recorded loops are delivered to their host pointers
 * by op2h_graph_sync(), as (host value at record time) op (device result), and
 * each launch also appends them to a device-side history read by
 * op2h_graph_history().  Returns 0 on success. */
int op2h_graph_begin(void);
int op2h_graph_end(void);
int op2h_graph_launch(int count);
int op2h_graph_sync(void);
int op2h_graph_history(int max_records, double *out, int *nrecords);
void op2h_graph_destroy(void);

/* Device-side timing of every kernel launch, per loop (HIP events on the
 * library's stream).  op2h_loop_stats fills, for loop `name`, the number of
 * timed calls and the summed kernel milliseconds; returns -1 if unknown. */
void op2h_enable_kernel_timing(int on);
int op2h_loop_stats(const char *name, int *calls, double *kernel_ms);
void op2h_reset_stats(void);

/* Library stream (hipStream_t) and a device synchronisation point. */
void *op2h_stream(void);
void op2h_device_sync(void);

/* Version string of the library. */
const char *op2h_version(void);

#ifdef __cplusplus
} /* extern "C" */

#include <cstddef>
#include <cstring>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <utility>  // op_seq.h:8, op_lib_core.h:65-68

/* ---- C++ templates (op_lib_cpp.h:127-203) ---- */
namespace op2h_detail {
inline bool type_ok(const double *, const char *t) {
  return !strcmp(t, "double") || !strcmp(t, "double:soa") || !strcmp(t, "r8");
}
inline bool type_ok(const float *, const char *t) {
  return !strcmp(t, "float") || !strcmp(t, "float:soa") || !strcmp(t, "r4");
}
inline bool type_ok(const int *, const char *t) {
  return !strcmp(t, "int") || !strcmp(t, "int:soa") || !strcmp(t, "i4");
}
inline bool type_ok(const bool *, const char *t) { return !strcmp(t, "bool"); }
}  // namespace op2h_detail

template <class T>
op_dat op_decl_dat(op_set set, int dim, char const *type, T *data,
                   char const *name) {
  if (!op2h_detail::type_ok(data, type)) {
    printf("incorrect type specified for dataset \"%s\" \n", name);
    exit(1);
  }
  return op_decl_dat_char(set, dim, type, sizeof(T), (char *)data, name);
}
template <class T>
void op_decl_const2(char const *name, int dim, char const *type, T *data) {
  if (!op2h_detail::type_ok(data, type)) {
    printf("incorrect type specified for constant \"%s\" \n", name);
    exit(1);
  }
  op_decl_const_char(dim, type, sizeof(T), (char *)data, name);
}
template <class T>
op_arg op_arg_gbl(T *data, int dim, char const *type, op_access acc) {
  return op_arg_gbl_char((char *)data, dim,
                         op2h_detail::type_ok(data, type) ? type : "error",
                         sizeof(T), acc);
}
template <class T>
op_arg op_opt_arg_gbl(int opt, T *data, int dim, char const *type,
                      op_access acc) {
  return op_opt_arg_gbl_char(opt, (char *)data, dim,
                             op2h_detail::type_ok(data, type) ? type : "error",
                             sizeof(T), acc);
}
template <class T>
op_arg op_arg_info(T *data, int dim, char const *type, int ref) {
  return op_arg_info_char((char *)data, dim,
                          op2h_detail::type_ok(data, type) ? type : "error",
                          sizeof(T), ref);
}
template <class T>
op_dat op_decl_dat_temp(op_set set, int dim, char const *type, T *,
                        char const *name) {
  return op_decl_dat_temp_char(set, dim, type, sizeof(T), name);
}
inline int op_free_dat_temp(op_dat dat) { return op_free_dat_temp_char(dat); }
template <class T> void op_fetch_data(op_dat dat, T *usr_ptr) {
  op_fetch_data_char(dat, (char *)usr_ptr);
}
template <class T>
void op_fetch_data_idx(op_dat dat, T *usr_ptr, int low, int high) {
  op_fetch_data_idx_char(dat, (char *)usr_ptr, low, high);
}
#endif /* __cplusplus */

#endif /* OP2_HIP_H */
