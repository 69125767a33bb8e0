// op2h_runtime.cpp — host side of the MI355X-native OP2 runtime.
//
// Registry, declarations and argument constructors keep the semantics of
// op2/src/core/op_lib_core.cpp (cited per function); the device side is this
// backend's own: one HIP stream, dats AoS or SoA with a stride padded to a
// multiple of 64 elements (one wave64 per 64 consecutive elements), plans
// uploaded once and cached per loop signature, gbl reductions delivered from
// device result slots, and optional HIP-graph capture of whole time steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "op2_hip_rt.h"

#include <rocprofiler-sdk-roctx/roctx.h>
#include "op2h_plan_build.h"
#include "op2h_profile.h"
#include "op2h_reduce.h"
#include "op2h_rcm.h"
#include "op2h_rcm_gpu.h"
#include "op2amd_costmodel.h"

// ---------------------------------------------------------------------------
// global state (op_lib_core.cpp:55-86)
// ---------------------------------------------------------------------------
int OP_diags = 0, OP_part_size = 0, OP_block_size = 256, OP_auto_soa = 0;
int OP_hybrid_gpu = 0, OP_cache_line_size = 128, OP_maps_base_index = 0;
int OP_kern_max = 0, OP_kern_curr = 0;
op_kernel *OP_kernels = nullptr;

#define HIPCHECK(x)                                                          \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace {

double wall_now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

char *copy_str(const char *s) {
  size_t n = std::strlen(s) + 1;
  char *p = (char *)std::malloc(n);
  std::memcpy(p, s, n);
  return p;
}

struct LoopConfig {
  int plan = OP2H_PLAN_AUTO;
  int block_size = 0;
};

struct TuneState {
  std::vector<std::pair<int, int>> cands;  // (plan, block)
  std::vector<double> best_ms;             // per candidate, min over trials
  size_t next = 0;
  int trials_done = 0;
  bool done = false;
  int plan = 0, block = 0;
};

// The cost model of op2h_set_plan_model and the loop descriptions it saw.
struct PlanModel {
  std::vector<op2amd_mlp *> members;
  std::vector<float> mean, scale;
  int top = 1;
};

struct ModelLoop {
  double set_size = 0, target_size = 0, arity = 1, dim = 1, is_double = 1, arith = 16,
         bandwidth = 0;
  int access = 0;  // 0 read, 1 inc, 2 rw
};

struct LoopStats {
  int calls = 0;
  double kernel_ms = 0.0;
};

struct PendingTimer {
  int kernel_idx;
  hipEvent_t a, b;
};

// A gbl result recorded inside a captured graph, delivered at graph sync.
struct DeferredGbl {
  op_access acc;
  int type;  // 0 double, 1 float, 2 int
  int dim;
  char *host_ptr;
  char *info_ptr;  // linked info arg for MIN/MAX (may be null)
  int info_size;
  std::vector<char> snapshot;       // host value at record time
  std::vector<char> info_snapshot;  // host info value at record time
  char *slot_h;                     // pinned copy written by the graph
  size_t bytes;                     // slot bytes (value + info)
};

struct Runtime {
  bool initialised = false;
  hipStream_t stream = nullptr;
  std::vector<op_set> sets;
  std::vector<op_map> maps;
  std::vector<op_dat> dats;
  std::map<std::string, LoopConfig> loop_cfg;
  LoopConfig all_cfg;
  std::map<int, LoopStats> stats;
  std::vector<PendingTimer> pending;
  std::vector<hipEvent_t> event_pool;
  bool kernel_timing = false;
  int sync_mode = OP2H_SYNC_EVERY_LOOP;
  int plan_on_device = 0;
  // plans
  std::vector<std::unique_ptr<op2h::StagedPlanRec>> staged;
  std::vector<std::unique_ptr<op2h::ColorPlanRec>> color;
  std::vector<std::unique_ptr<op2h::OwnerPlanRec>> owner;
  std::map<op_map, std::unique_ptr<op2h::GatherPlanRec>> gather;
  // reductions
  void *red_scratch = nullptr;
  size_t red_scratch_bytes = 0;
  unsigned int *red_counter = nullptr;
  // gbl slots: persistent per (kernel, arg) outside capture
  std::map<std::pair<int, int>, std::pair<void *, void *>> slots;
  std::vector<std::pair<void *, void *>> capture_slots;  // (dev, pinned) pool
  int capture_slot_next = 0;
  // graph capture
  bool capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<DeferredGbl> deferred;
  std::vector<std::pair<void *, size_t>> history_sources;  // device slots
  // deferred reductions: slot -> (kind 1 sum / 2 max / 3 min, nparts, vals, locs)
  struct Partials {
    int kind, nparts;
    const double *vals;
    const int *locs;
  };
  std::map<void *, Partials> partial_of_slot;
  char *partial_pool = nullptr;
  size_t partial_pool_bytes = 0, partial_pool_next = 0;
  double *history = nullptr;
  unsigned long long *history_count = nullptr;
  unsigned int *history_done = nullptr;  // history kernel's workgroup ticket
  int history_words = 0;   // words per launch
  int history_capacity = 0;  // launches
  std::map<op_map, long long> distinct_targets;  // per map, for byte counts
  std::vector<std::vector<int>> set_new_of_old;  // composed op_renumber perms
  // transparent renumbering (OP_HIP_RENUMBER): before the first loop the
  // sets are renumbered as op_renumber does, but the application keeps its
  // own numbering: fetches, dumps and element indices are translated
  bool transparent = false, transparent_done = false;
  // under transparent renumbering: each map's table in the device numbering,
  // by map index (the op_map's own table keeps the application's)
  std::vector<std::vector<int>> dev_maps;
  std::string transparent_map;                    // base map named by the token
  std::vector<std::vector<int>> set_old_of_new;   // per renumbered set
  std::vector<int *> old_index_d;                 // device copies of those
  std::map<std::string, std::pair<void *, size_t>> user_plans;
  // autotuning
  bool autotune = false;
  int tune_trials = 2;
  std::map<std::string, TuneState> tune;
  PlanModel model;                          // kept across op_init/op_exit
  std::map<std::string, int> loop_arith;    // op2h_set_loop_arith
  std::map<std::string, ModelLoop> model_loops;
  std::vector<op2h_tuning_record> tune_records;
  hipEvent_t tune_ev[2] = {nullptr, nullptr};
  // a kernel launch a stub left pending for the next stub to fuse
  // (op2h_defer); flushed by everything that is not that stub
  op2h_flush_fn pend_fn = nullptr;
  void *pend_arg = nullptr;
  std::string pend_tag, accept_tag;
  bool fusion = true;  // op2h_set_fusion; kept across op_init/op_exit
  int streaming = -1;  // op2h_set_streaming: -1 by size, 0 never, 1 always
  unsigned plan_epoch = 1;  // advanced whenever the plans are dropped
  unsigned long long fused = 0;  // pending launches taken over by a later stub
  // the current loop's node in the op_profile tree (loops do not nest)
  bool prof_loop = false;
  double prof_t0 = 0.0, prof_comp_t0 = 0.0, prof_kms0 = 0.0;
  bool timing_on() const { return kernel_timing || (prof_loop && op2h::profile_level() >= 3); }
};

Runtime &rt() {
  static Runtime R;
  return R;
}

const char *app_rows(op_dat dat, std::vector<char> &buf);
void transparent_renumber();

// The device numbering of a set under transparent renumbering (new_of_old),
// or null when the set keeps the application's.
const int *device_number(op_set set) {
  Runtime &R = rt();
  if (!R.transparent || !R.transparent_done || (size_t)set->index >= R.set_new_of_old.size() ||
      R.set_new_of_old[set->index].empty())
    return nullptr;
  return R.set_new_of_old[set->index].data();
}

void ensure_init() {
  Runtime &R = rt();
  if (R.initialised) return;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    printf("op2hip: no HIP device available; this backend has no CPU "
           "fallback\n");
    exit(-1);
  }
  HIPCHECK(hipSetDevice(0));
  HIPCHECK(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
  HIPCHECK(hipMalloc(&R.red_counter, 64 * sizeof(unsigned int)));
  HIPCHECK(hipMemsetAsync(R.red_counter, 0, 64 * sizeof(unsigned int), R.stream));
  HIPCHECK(hipStreamSynchronize(R.stream));
  OP_hybrid_gpu = 1;
  R.initialised = true;
}

int type_code(const char *t) {
  if (!std::strncmp(t, "double", 6) || !std::strcmp(t, "r8")) return 0;
  if (!std::strncmp(t, "float", 5) || !std::strcmp(t, "r4")) return 1;
  if (!std::strncmp(t, "int", 3) || !std::strcmp(t, "i4")) return 2;
  if (!std::strcmp(t, "bool")) return 3;
  return -1;
}

int elem_bytes_of(const char *t, int fallback) {
  switch (type_code(t)) {
    case 0: return 8;
    case 1: return 4;
    case 2: return 4;
    case 3: return 1;
  }
  return fallback;
}

bool dat_is_soa(op_dat d) {
  if (d->dim <= 1) return false;
  if (std::strstr(d->type, ":soa")) return true;
  return OP_auto_soa != 0;
}

int round64(int n) { return (n + 63) & ~63; }

// --- layout transforms (op_cuda_decl.cpp:71-99, op_cuda_rt_support.cpp:373-396) ---
void upload_dat(op_dat d) {
  Runtime &R = rt();
  op_set s = d->set;
  size_t n = (size_t)s->size;
  size_t bytes = (size_t)d->size * (size_t)std::max(s->stride, 1);
  if (!d->data_d) {
    HIPCHECK(hipMalloc(&d->data_d, std::max(bytes, (size_t)64)));
    // on the library's stream: it does not synchronise with the null stream,
    // so a null-stream memset could land after the copy below
    HIPCHECK(hipMemsetAsync(d->data_d, 0, std::max(bytes, (size_t)64), R.stream));
  }
  // the host rows are in the application's numbering; under transparent
  // renumbering row i goes to the device's row p[i]
  const int *p = device_number(s);
  std::vector<char> tmp;
  if (!dat_is_soa(d)) {
    const char *src = d->data;
    if (p) {
      tmp.resize((size_t)d->size * n);
      for (size_t i = 0; i < n; ++i)
        std::memcpy(&tmp[(size_t)p[i] * d->size], d->data + i * d->size, d->size);
      src = tmp.data();
    }
    HIPCHECK(hipMemcpyAsync(d->data_d, src, (size_t)d->size * n, hipMemcpyHostToDevice,
                            R.stream));
  } else {
    int eb = d->size / d->dim;
    tmp.resize(bytes);
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < d->dim; ++k)
        std::memcpy(&tmp[((size_t)k * s->stride + (p ? (size_t)p[i] : i)) * eb],
                    d->data + (i * d->dim + k) * eb, eb);
    HIPCHECK(hipMemcpyAsync(d->data_d, tmp.data(), bytes, hipMemcpyHostToDevice,
                            R.stream));
  }
  HIPCHECK(hipStreamSynchronize(R.stream));
  d->dirty_hd = 0;
}

void download_dat(op_dat d) {
  Runtime &R = rt();
  if (!d->data_d) return;
  op_set s = d->set;
  size_t n = (size_t)s->size;
  const int *p = device_number(s);  // device row of the application's row i
  if (!dat_is_soa(d) && !p) {
    HIPCHECK(hipMemcpyAsync(d->data, d->data_d, (size_t)d->size * n,
                            hipMemcpyDeviceToHost, R.stream));
    HIPCHECK(hipStreamSynchronize(R.stream));
  } else if (!dat_is_soa(d)) {
    std::vector<char> tmp((size_t)d->size * n);
    HIPCHECK(hipMemcpyAsync(tmp.data(), d->data_d, tmp.size(), hipMemcpyDeviceToHost,
                            R.stream));
    HIPCHECK(hipStreamSynchronize(R.stream));
    for (size_t i = 0; i < n; ++i)
      std::memcpy(d->data + i * d->size, &tmp[(size_t)p[i] * d->size], d->size);
  } else {
    int eb = d->size / d->dim;
    size_t bytes = (size_t)d->size * s->stride;
    std::vector<char> tmp(bytes);
    HIPCHECK(hipMemcpyAsync(tmp.data(), d->data_d, bytes, hipMemcpyDeviceToHost,
                            R.stream));
    HIPCHECK(hipStreamSynchronize(R.stream));
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < d->dim; ++k)
        std::memcpy(d->data + (i * d->dim + k) * eb,
                    &tmp[((size_t)k * s->stride + (p ? (size_t)p[i] : i)) * eb], eb);
  }
  d->dirty_hd = 0;
}

hipEvent_t take_event() {
  Runtime &R = rt();
  if (!R.event_pool.empty()) {
    hipEvent_t e = R.event_pool.back();
    R.event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCHECK(hipEventCreate(&e));
  return e;
}

void resolve_timers() {
  Runtime &R = rt();
  if (R.pending.empty()) return;
  HIPCHECK(hipStreamSynchronize(R.stream));
  for (auto &p : R.pending) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, p.a, p.b));
    R.stats[p.kernel_idx].kernel_ms += ms;
    R.stats[p.kernel_idx].calls += 1;
    R.event_pool.push_back(p.a);
    R.event_pool.push_back(p.b);
  }
  R.pending.clear();
}

// Combine a device result `res` into host value `host` (dim elements).
template <class T>
bool combine(op_access acc, T *host, const T *res, int dim, bool *improved) {
  for (int d = 0; d < dim; ++d) {
    if (acc == OP_INC) {
      host[d] += res[d];
    } else if (acc == OP_MAX) {
      if (res[d] > host[d]) {
        host[d] = res[d];
        if (improved) improved[d] = true;
      }
    } else if (acc == OP_MIN) {
      if (res[d] < host[d]) {
        host[d] = res[d];
        if (improved) improved[d] = true;
      }
    } else if (acc == OP_WRITE || acc == OP_RW) {
      host[d] = res[d];
    }
  }
  return true;
}

void combine_gbl(op_access acc, int type, int dim, char *host, const char *res,
                 char *info_host, const char *info_res, int info_size) {
  bool improved[64] = {false};
  int dd = std::min(dim, 64);
  switch (type) {
    case 0: combine(acc, (double *)host, (const double *)res, dd, improved); break;
    case 1: combine(acc, (float *)host, (const float *)res, dd, improved); break;
    case 2: combine(acc, (int *)host, (const int *)res, dd, improved); break;
    default: break;
  }
  if (info_host && info_res) {
    int per = info_size / std::max(dim, 1);
    for (int d = 0; d < dd; ++d)
      if (improved[d]) std::memcpy(info_host + d * per, info_res + d * per, per);
  }
}

}  // namespace

// ===========================================================================
// lifecycle and declarations
// ===========================================================================
extern "C" void op_init(int argc, char **argv, int diags) {
  OP_diags = diags;
  // argv tokens as op_set_args (op_lib_core.cpp:205-281)
  for (int n = 1; n < argc; ++n) {
    const char *a = argv[n];
    if (!std::strncmp(a, "OP_BLOCK_SIZE=", 14)) OP_block_size = atoi(a + 14);
    else if (!std::strncmp(a, "OP_PART_SIZE=", 13)) OP_part_size = atoi(a + 13);
    else if (!std::strncmp(a, "OP_CACHE_LINE_SIZE=", 19)) OP_cache_line_size = atoi(a + 19);
    else if (!std::strcmp(a, "OP_AUTO_SOA")) OP_auto_soa = 1;
    else if (!std::strncmp(a, "OP_MAPS_BASE_INDEX=", 19)) OP_maps_base_index = atoi(a + 19);
    else if (!std::strcmp(a, "OP_HIP_KERNEL_TIMING")) op2h_enable_kernel_timing(1);
    else if (!std::strncmp(a, "OP_HIP_STREAMING=", 17))
      op2h_set_streaming(!std::strcmp(a + 17, "on") ? 1 : !std::strcmp(a + 17, "off") ? 0 : -1);
    else if (!std::strcmp(a, "OP_HIP_RENUMBER") || !std::strncmp(a, "OP_HIP_RENUMBER=", 16)) {
      rt().transparent = true;
      rt().transparent_map = a[15] == '=' ? std::string(a + 16) : std::string();
    }
    else if (!std::strncmp(a, "OP_HIP_PLAN=", 12)) {
      // OP_HIP_PLAN=<loop>:<plan>[:<block>]
      std::string s(a + 12);
      size_t c1 = s.find(':');
      if (c1 == std::string::npos) continue;
      std::string loop = s.substr(0, c1), rest = s.substr(c1 + 1);
      size_t c2 = rest.find(':');
      std::string plan = rest.substr(0, c2);
      int bs = c2 == std::string::npos ? 0 : atoi(rest.c_str() + c2 + 1);
      int p = plan == "atomic" ? OP2H_PLAN_ATOMIC
              : plan == "color" ? OP2H_PLAN_COLOR
              : plan == "staged" ? OP2H_PLAN_STAGED
              : plan == "gather" ? OP2H_PLAN_GATHER
              : plan == "owner" ? OP2H_PLAN_OWNER : OP2H_PLAN_AUTO;
      op2h_set_loop_plan(loop == "*" ? nullptr : loop.c_str(), p, bs);
    }
  }
  ensure_init();
}

extern "C" void op_init_soa(int argc, char **argv, int diags, int soa) {
  OP_auto_soa = soa;
  op_init(argc, argv, diags);
}

extern "C" void op_exit(void) {
  Runtime &R = rt();
  if (!R.initialised) return;
  op2h_flush_deferred();
  resolve_timers();
  HIPCHECK(hipStreamSynchronize(R.stream));
  op2h_graph_destroy();
  R.staged.clear();
  ++R.plan_epoch;
  R.color.clear();
  R.owner.clear();
  R.gather.clear();
  for (auto &kv : R.user_plans) (void)hipFree(kv.second.first);
  R.user_plans.clear();
  for (op_dat d : R.dats) {
    if (!d) continue;
    if (d->data_d) (void)hipFree(d->data_d);
    if (!d->user_managed) std::free(d->data);
    std::free((void *)d->name);
    std::free((void *)d->type);
    std::free(d);
  }
  for (op_map m : R.maps) {
    if (m->map_d) (void)hipFree(m->map_d);
    std::free(m->map);
    std::free((void *)m->name);
    std::free(m);
  }
  for (op_set s : R.sets) {
    std::free((void *)s->name);
    std::free(s);
  }
  for (auto &kv : R.slots) {
    (void)hipFree(kv.second.first);
    (void)hipHostFree(kv.second.second);
  }
  if (R.red_scratch) (void)hipFree(R.red_scratch);
  if (R.red_counter) (void)hipFree(R.red_counter);
  for (auto e : R.event_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(R.stream);
  for (int k = 0; k < OP_kern_max; ++k) std::free(OP_kernels[k].times);
  std::free(OP_kernels);
  OP_kernels = nullptr;
  OP_kern_max = 0;
  R.sets.clear();
  R.maps.clear();
  R.dats.clear();
  R.slots.clear();
  R.capture_slots.clear();
  R.stats.clear();
  R.distinct_targets.clear();
  R.set_new_of_old.clear();
  for (int *p : R.old_index_d)
    if (p) (void)hipFree(p);
  R.old_index_d.clear();
  R.set_old_of_new.clear();
  R.dev_maps.clear();
  R.transparent = R.transparent_done = false;
  R.transparent_map.clear();
  R.tune.clear();
  R.autotune = false;
  if (R.tune_ev[0]) (void)hipEventDestroy(R.tune_ev[0]);
  if (R.tune_ev[1]) (void)hipEventDestroy(R.tune_ev[1]);
  R.tune_ev[0] = R.tune_ev[1] = nullptr;
  R.loop_cfg.clear();
  R.all_cfg = LoopConfig();
  R.event_pool.clear();
  R.red_scratch = nullptr;
  R.red_scratch_bytes = 0;
  R.red_counter = nullptr;
  R.stream = nullptr;
  R.initialised = false;
}

extern "C" op_set op_decl_set(idx_l_t size, char const *name) {
  // op_lib_core.cpp:391-418
  ensure_init();
  if (size < 0) {
    printf(" op_decl_set error -- negative/zero size for set: %s\n", name);
    exit(-1);
  }
  op_set s = (op_set)std::calloc(1, sizeof(op_set_core));
  s->index = (int)rt().sets.size();
  s->size = size;
  s->core_size = size;
  s->name = copy_str(name);
  s->stride = round64(std::max(size, 1));
  rt().sets.push_back(s);
  return s;
}

extern "C" op_map op_decl_map(op_set from, op_set to, int dim, idx_l_t *imap,
                              char const *name) {
  // op_lib_core.cpp:419-500 (copies the table; optional 1-based conversion)
  ensure_init();
  if (!from) {
    printf(" op_decl_map error -- invalid 'from' set for map %s\n", name);
    exit(-1);
  }
  if (!to) {
    printf("op_decl_map error -- invalid 'to' set for map %s\n", name);
    exit(-1);
  }
  if (dim <= 0) {
    printf("op_decl_map error -- negative/zero dimension for map %s\n", name);
    exit(-1);
  }
  size_t n = (size_t)from->size * dim;
  op_map m = (op_map)std::calloc(1, sizeof(op_map_core));
  m->index = (int)rt().maps.size();
  m->from = from;
  m->to = to;
  m->dim = dim;
  m->name = copy_str(name);
  m->map = (int *)std::malloc(std::max(n, (size_t)1) * sizeof(int));
  if (imap) std::memcpy(m->map, imap, n * sizeof(int));
  else std::memset(m->map, 0, n * sizeof(int));
  if (OP_maps_base_index == 1 && imap)
    for (size_t i = 0; i < n; ++i) m->map[i]--;
  // a map entry outside the target set would make every kernel that reads
  // through it access memory out of bounds: refuse it here
  for (size_t i = 0; i < n; ++i)
    if (m->map[i] < 0 || m->map[i] >= to->size) {
      printf("op_decl_map error -- invalid data for map %s\n", name);
      printf("element = %zu, dimension = %zu, map = %d\n", i / dim, i % dim,
             m->map[i]);
      exit(-1);
    }
  // declared after transparent renumbering: the device's table goes from
  // the application's numbering to the device's (entries through the target
  // set's, rows through the source set's); the map's own table stays as given
  Runtime &R = rt();
  if (R.dev_maps.size() <= (size_t)m->index) R.dev_maps.resize(m->index + 1);
  const int *pt = device_number(to), *pf = device_number(from);
  if (pt || pf) {
    std::vector<int> ent(m->map, m->map + n);
    if (pt)
      for (size_t i = 0; i < n; ++i) ent[i] = pt[ent[i]];
    std::vector<int> &dev = R.dev_maps[m->index];
    if (pf) {
      dev.resize(n);
      for (int i = 0; i < from->size; ++i)
        std::memcpy(&dev[(size_t)pf[i] * dim], &ent[(size_t)i * dim], dim * sizeof(int));
    } else {
      dev.swap(ent);
    }
  }
  HIPCHECK(hipMalloc(&m->map_d, std::max(n, (size_t)1) * sizeof(int)));
  HIPCHECK(hipMemcpy(m->map_d, op2h::host_map(m), n * sizeof(int), hipMemcpyHostToDevice));
  R.maps.push_back(m);
  return m;
}

extern "C" op_map op_decl_map_long(op_set from, op_set to, int dim,
                                   idx_g_t *imap, char const *name) {
  std::vector<int> tmp((size_t)from->size * dim);
  for (size_t i = 0; i < tmp.size(); ++i) tmp[i] = (int)imap[i];
  return op_decl_map(from, to, dim, tmp.data(), name);
}

namespace {
// op_lib_core.cpp:502-576 (always copies: OP_NO_REALLOC is not supported);
// `temp`: zeroed on the device directly, nothing uploaded
op_dat decl_dat(op_set set, int dim, char const *type, int size, char *data,
                char const *name, bool temp) {
  ensure_init();
  if (!set) {
    printf("op_decl_dat error -- invalid set for data: %s\n", name);
    exit(-1);
  }
  if (dim <= 0) {
    printf("op_decl_dat error -- negative/zero dimension for data: %s\n", name);
    exit(-1);
  }
  op_dat d = (op_dat)std::calloc(1, sizeof(op_dat_core));
  d->index = (int)rt().dats.size();
  d->set = set;
  d->dim = dim;
  d->size = dim * size;
  size_t bytes = (size_t)d->size * set->size;
  d->data = (char *)std::calloc(std::max(bytes, (size_t)1), 1);
  // the application's rows (uploads move them to the device numbering)
  if (data) std::memcpy(d->data, data, bytes);
  d->name = copy_str(name);
  d->type = copy_str(type);
  d->dirtybit = 1;
  rt().dats.push_back(d);
  if (!temp) {
    upload_dat(d);
  } else {
    Runtime &R = rt();
    const size_t dbytes =
        std::max((size_t)d->size * (size_t)std::max(set->stride, 1), (size_t)64);
    HIPCHECK(hipMalloc(&d->data_d, dbytes));
    HIPCHECK(hipMemsetAsync(d->data_d, 0, dbytes, R.stream));
    HIPCHECK(hipStreamSynchronize(R.stream));
    d->dirty_hd = 0;
  }
  return d;
}
}  // namespace

extern "C" op_dat op_decl_dat_char(op_set set, int dim, char const *type,
                                   int size, char *data, char const *name) {
  return decl_dat(set, dim, type, size, data, name, false);
}

// op_lib_core.cpp:643-680: a temporary dat may not share its name (with
// its set, dim, type and element size) with a live dat; it starts zeroed
// (op_decl_dat_core's calloc), here on the device without a host copy.
extern "C" op_dat op_decl_dat_temp_char(op_set set, int dim, char const *type,
                                        int size, char const *name) {
  ensure_init();
  for (op_dat d : rt().dats)
    if (d && d->set == set && d->dim == dim && d->size == dim * size &&
        !std::strcmp(d->name, name) && !std::strcmp(d->type, type)) {
      printf("op_dat with name %s already exists, cannot create temporary op_dat\n ", name);
      exit(2);
    }
  return decl_dat(set, dim, type, size, nullptr, name, true);
}

// 1 when the dat was found and freed, -1 otherwise (op_lib_core.cpp:657-680)
extern "C" int op_free_dat_temp_char(op_dat dat) {
  Runtime &R = rt();
  for (auto &d : R.dats)
    if (d && d == dat) {
      op2h_flush_deferred();
      HIPCHECK(hipStreamSynchronize(R.stream));
      if (d->data_d) (void)hipFree(d->data_d);
      std::free(d->data);
      std::free((void *)d->name);
      std::free((void *)d->type);
      std::free(d);
      d = nullptr;
      return 1;
    }
  return -1;
}

// ===========================================================================
// argument constructors (op_lib_core.cpp:832-1060)
// ===========================================================================
namespace {
const char *canon_type(const char *typ) {
  if (!std::strcmp(typ, "double") || !std::strcmp(typ, "r8") ||
      !std::strcmp(typ, "real*8") || !std::strcmp(typ, "real(8)"))
    return "double";
  if (!std::strcmp(typ, "float") || !std::strcmp(typ, "r4") ||
      !std::strcmp(typ, "real*4") || !std::strcmp(typ, "real(4)"))
    return "float";
  if (!std::strcmp(typ, "int") || !std::strcmp(typ, "i4") ||
      !std::strcmp(typ, "integer*4") || !std::strcmp(typ, "integer(4)"))
    return "int";
  if (!std::strcmp(typ, "bool")) return "bool";
  return typ;
}
}  // namespace

extern "C" op_arg op_opt_arg_dat(int opt, op_dat dat, int idx, op_map map,
                                 int dim, char const *typ, op_access acc) {
  op_arg a;
  std::memset(&a, 0, sizeof(a));
  a.index = -1;
  a.opt = opt;
  a.argtype = OP_ARG_DAT;
  a.dat = dat;
  a.map = map;
  a.dim = dim;
  a.idx = idx;
  if (dat) {
    a.size = dat->size;
    a.data = dat->data;
    a.data_d = dat->data_d;
  } else {
    a.size = -1;
  }
  a.map_data = map ? map->map : nullptr;
  a.map_data_d = map ? map->map_d : nullptr;
  a.type = canon_type(typ);
  a.acc = acc;
  return a;
}

extern "C" op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim,
                             char const *typ, op_access acc) {
  op_arg a = op_opt_arg_dat(1, dat, idx, map, dim, typ, acc);
  if (idx == -1) {
    a.map_data = nullptr;
    a.map_data_d = nullptr;
  }
  return a;
}

extern "C" op_arg op_opt_arg_gbl_char(int opt, char *data, int dim,
                                      const char *typ, int size,
                                      op_access acc) {
  op_arg a;
  std::memset(&a, 0, sizeof(a));
  a.argtype = OP_ARG_GBL;
  a.opt = opt;
  a.dim = dim;
  a.idx = -1;
  a.size = dim * size;
  a.data = data;
  a.type = canon_type(typ);
  a.acc = acc;
  a.index = -1;
  if (!data) a.opt = 0;
  return a;
}

extern "C" op_arg op_arg_gbl_char(char *data, int dim, const char *typ,
                                  int size, op_access acc) {
  return op_opt_arg_gbl_char(1, data, dim, typ, size, acc);
}

extern "C" op_arg op_arg_info_char(char *data, int dim, const char *typ,
                                   int size, int ref) {
  op_arg a = op_opt_arg_gbl_char(1, data, dim, typ, size, ref);
  a.argtype = OP_ARG_INFO;
  return a;
}

extern "C" op_arg op_arg_idx(int idx, op_map map) {
  op_arg a;
  std::memset(&a, 0, sizeof(a));
  a.index = -1;
  a.opt = 1;
  a.argtype = OP_ARG_IDX;
  a.map = map;
  a.idx = idx;
  a.dim = 1;
  a.size = sizeof(int);
  a.type = "int";
  a.acc = OP_READ;
  a.map_data = (idx == -1 || !map) ? nullptr : map->map;
  a.map_data_d = (idx == -1 || !map) ? nullptr : map->map_d;
  return a;
}

extern "C" void op_arg_check(op_set set, int m, op_arg arg, int *ninds,
                             const char *name) {
  // op_lib_core.cpp:754-829: prints, never exits
  auto err = [&](const char *msg) {
    printf("error: arg %d in kernel \"%s\"\n", m, name);
    printf("%s \n", msg);
  };
  if (arg.opt == 0) return;
  if (arg.argtype == OP_ARG_DAT) {
    if (!set) err("invalid set");
    if (!arg.map && arg.dat->set != set)
      printf("dataset dat %s on set %s does not match loop set %s\n",
             arg.dat->name, arg.dat->set->name, set ? set->name : "?");
    if (arg.map) {
      if (arg.map->from != set) {
        err("mapping error");
        printf("map from set %s does not match set \n", arg.map->from->name);
      }
      if (arg.map->to != arg.dat->set) {
        err("mapping error");
        printf("map %s to set %s does not match dat %s set %s\n", arg.map->name,
               arg.map->to->name, arg.dat->name, arg.dat->set->name);
      }
    }
    if ((!arg.map && arg.idx != -1) ||
        (arg.map && (arg.idx >= arg.map->dim || arg.idx < -1 * arg.map->dim)))
      err("invalid index");
    if (arg.dat->dim != arg.dim) err("dataset dim does not match declared dim");
    if (std::strcmp(canon_type(arg.dat->type), arg.type) &&
        std::strncmp(arg.dat->type, arg.type, std::strlen(arg.type)))
      err("dataset type does not match declared type");
    if (arg.idx >= 0) (*ninds)++;
  }
  if (arg.argtype == OP_ARG_GBL || arg.argtype == OP_ARG_INFO) {
    if (!std::strcmp(arg.type, "error")) err("datatype does not match declared type");
    if (arg.dim < 0) err("dimension should be positive");
    if (!arg.data) err("NULL pointer for global data");
  }
  if (arg.argtype == OP_ARG_IDX) {
    if (arg.idx == -1 && arg.map) err("op_arg_idx with -1 index but non-OP_ID map");
    if (arg.idx != -1 && !arg.map) err("op_arg_idx map index not -1, but no valid map given");
  }
}

// ===========================================================================
// data access (op_cuda_rt_support.cpp:276-307,373-396; op_lib_core.cpp:1283-1369)
// ===========================================================================
extern "C" void op_upload_dat(op_dat dat) { upload_dat(dat); }
extern "C" void op_download_dat(op_dat dat) { download_dat(dat); }
extern "C" void op_upload_all(void) {
  for (op_dat d : rt().dats)
    if (d && d->dirty_hd == 1) upload_dat(d);
}

extern "C" void op_fetch_data_char(op_dat dat, char *usr_ptr) {
  if (rt().capturing) {
    printf("op_fetch_data error -- not allowed during graph capture\n");
    exit(-1);
  }
  op2h_flush_deferred();
  if (dat->dirty_hd == 2) download_dat(dat);
  std::vector<char> buf;
  std::memcpy(usr_ptr, app_rows(dat, buf), (size_t)dat->size * dat->set->size);
}

extern "C" void op_fetch_data_idx_char(op_dat dat, char *usr_ptr, int low,
                                       int high) {
  if (low < 0 || high >= dat->set->size || low > high) {
    printf("op_fetch_data_idx error -- range %d..%d outside set %s\n", low, high,
           dat->set->name);
    exit(-1);
  }
  op2h_flush_deferred();
  if (dat->dirty_hd == 2) download_dat(dat);
  std::vector<char> buf;
  std::memcpy(usr_ptr, app_rows(dat, buf) + (size_t)low * dat->size,
              (size_t)(high - low + 1) * dat->size);
}

extern "C" idx_g_t op_get_size(op_set set) { return set->size; }

extern "C" void op_print_dat_to_binfile(op_dat dat, const char *file_name) {
  // op_lib_core.cpp:1283-1307: int count, int dim, then raw AoS data
  op2h_flush_deferred();
  if (dat->dirty_hd == 2) download_dat(dat);
  FILE *fp = std::fopen(file_name, "wb");
  if (!fp) {
    printf("can't open file %s\n", file_name);
    exit(2);
  }
  int count = dat->set->size, dim = dat->dim;
  std::vector<char> buf;
  const char *rows = app_rows(dat, buf);
  if (std::fwrite(&count, sizeof(int), 1, fp) < 1 ||
      std::fwrite(&dim, sizeof(int), 1, fp) < 1 ||
      std::fwrite(rows, dat->size, count, fp) < (size_t)count) {
    printf("error writing to %s\n", file_name);
    exit(2);
  }
  std::fclose(fp);
}

extern "C" void op_print_dat_to_txtfile(op_dat dat, const char *file_name) {
  // op_lib_core.cpp:1309-1369
  op2h_flush_deferred();
  if (dat->dirty_hd == 2) download_dat(dat);
  FILE *fp = std::fopen(file_name, "w");
  if (!fp) {
    printf("can't open file %s\n", file_name);
    exit(2);
  }
  std::fprintf(fp, "%d %d\n", dat->set->size, dat->dim);
  int tc = type_code(dat->type);
  std::vector<char> buf;
  const char *rows = app_rows(dat, buf);
  for (int i = 0; i < dat->set->size; i++) {
    for (int j = 0; j < dat->dim; j++) {
      size_t k = (size_t)i * dat->dim + j;
      if (tc == 0) {
        double v = ((const double *)rows)[k];
        if (v == -0.0) v = +0.0;
        std::fprintf(fp, " %+2.15lE", v);
      } else if (tc == 1) {
        std::fprintf(fp, " %+f", ((const float *)rows)[k]);
      } else if (tc == 2) {
        std::fprintf(fp, " %+d", ((const int *)rows)[k]);
      } else {
        printf("Unknown type %s, cannot be written to file %s\n", dat->type,
               file_name);
        exit(2);
      }
    }
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}

// ===========================================================================
// timing and diagnostics (op_lib_core.cpp:1067-1167)
// ===========================================================================
extern "C" void op_timers_core(double *cpu, double *et) {
  *cpu = (double)clock() / CLOCKS_PER_SEC;
  *et = wall_now();
}
extern "C" void op_timers(double *cpu, double *et) {
  op2h_device_sync();
  op_timers_core(cpu, et);
}

extern "C" void op_timing_realloc(int kernel) {
  if (kernel < OP_kern_max) return;
  int old = OP_kern_max;
  OP_kern_max = kernel + 10;
  OP_kernels = (op_kernel *)std::realloc(OP_kernels, OP_kern_max * sizeof(op_kernel));
  for (int n = old; n < OP_kern_max; ++n) {
    std::memset(&OP_kernels[n], 0, sizeof(op_kernel));
    OP_kernels[n].name = "unused";
    OP_kernels[n].times = (double *)std::calloc(1, sizeof(double));
    OP_kernels[n].ntimes = 1;
  }
}

extern "C" void op_timing_output(void) {
  resolve_timers();
  if (OP_kern_max == 0) return;
  op_printf("\n  count   plan time     time(host)   kernel ms      GB/s      "
            "kernel name \n");
  op_printf(" -------------------------------------------------------------"
            "-------------------\n");
  for (int n = 0; n < OP_kern_max; ++n) {
    if (OP_kernels[n].count == 0) continue;
    auto it = rt().stats.find(n);
    double kms = it == rt().stats.end() ? 0.0 : it->second.kernel_ms;
    double secs = kms > 0 ? kms * 1e-3 : OP_kernels[n].time;
    double gbs = secs > 0 ? OP_kernels[n].transfer / secs * 1e-9 : 0.0;
    op_printf(" %6d;  %8.4f;  %10.4f;  %10.4f;  %8.2f;   %s \n", OP_kernels[n].count,
              OP_kernels[n].plan_time, OP_kernels[n].time, kms, gbs,
              OP_kernels[n].name);
  }
}

extern "C" void op_printf(const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  vprintf(format, ap);
  va_end(ap);
}

extern "C" void op_diagnostic_output(void) {
  if (OP_diags <= 1) return;
  printf("\n  OP diagnostic output\n");
  printf("  --------------------\n");
  printf("\n       set       size\n");
  for (op_set s : rt().sets) printf("%10s %10d\n", s->name, s->size);
  printf("\n       map        dim       from         to\n");
  for (op_map m : rt().maps)
    printf("%10s %10d %10s %10s\n", m->name, m->dim, m->from->name, m->to->name);
  printf("\n       dat        dim        set     layout\n");
  for (op_dat d : rt().dats)
    if (d)
      printf("%10s %10d %10s %10s\n", d->name, d->dim, d->set->name,
             dat_is_soa(d) ? "soa" : "aos");
  printf("\n");
}

extern "C" int op_is_root(void) { return 1; }

// ---- host memory with OP2_ALIGNMENT (op_lib_core.cpp:1388-1420) ----
extern "C" void *op_malloc(size_t size) {
  if (size == 0) return malloc(0);
  const size_t a = OP2_ALIGNMENT;
  return aligned_alloc(a, (size + a - 1) / a * a);
}
extern "C" void *op_calloc(size_t num, size_t size) {
  const size_t bytes = num * size;
  if (size != 0 && bytes / size != num) return nullptr;  // overflow
  void *p = op_malloc(bytes);
  if (p) std::memset(p, 0, bytes);
  return p;
}
extern "C" void *op_realloc(void *ptr, size_t size) {
  void *p = realloc(ptr, size);
  if (!p || ((uintptr_t)p % OP2_ALIGNMENT) == 0) return p;
  void *q = op_malloc(size);  // realloc moved it off the alignment
  if (q) std::memcpy(q, p, size);
  free(p);
  return q;
}
extern "C" void op_free(void *ptr) { free(ptr); }
extern "C" void op_rank(int *rank) { *rank = 0; }

// ---- single-rank no-ops (op_dummy_singlenode.cpp:43-208) ----
extern "C" void op_partition(const char *, const char *, op_set, op_map, op_dat) {}
extern "C" int op_mpi_halo_exchanges(op_set set, int, op_arg *) { return set->size; }
extern "C" int op_mpi_halo_exchanges_cuda(op_set set, int, op_arg *) { return set->size; }
extern "C" int op_mpi_halo_exchanges_grouped(op_set set, int, op_arg *, int) {
  return set->size;
}
extern "C" void op_mpi_set_dirtybit(int, op_arg *) {}
extern "C" void op_mpi_set_dirtybit_cuda(int, op_arg *) {}
extern "C" void op_mpi_wait_all(int, op_arg *) {}
extern "C" void op_mpi_wait_all_cuda(int, op_arg *) {}
extern "C" void op_mpi_wait_all_grouped(int, op_arg *, int) {}
extern "C" void op_mpi_reduce_float(op_arg *, float *) {}
extern "C" void op_mpi_reduce_double(op_arg *, double *) {}
extern "C" void op_mpi_reduce_int(op_arg *, int *) {}
extern "C" void op_mpi_reduce_bool(op_arg *, bool *) {}
extern "C" void op_mpi_barrier(void) {}

// ===========================================================================
// renumbering (op2/src/externlib/op_renumber.cpp:72-198,224-505)
// ===========================================================================
namespace {

struct Renumber {
  std::vector<std::vector<int>> perm, iperm;  // per set: new_of_old, old_of_new

  // op_renumber.cpp:72-151: give `to` an order derived from `from`, then
  // recurse over the maps touching `to`, in declaration order
  void propagate(op_set from, op_set to) {
    Runtime &R = rt();
    if (to->size == 0) return;
    std::vector<int> &pt = perm[to->index];
    if (pt.empty()) {  // a map to->from: stable sort by new index of its first column
      for (op_map m : R.maps)
        if (m->to == from && m->from == to) {
          std::vector<int> order(to->size);
          for (int i = 0; i < to->size; ++i) order[i] = i;
          const std::vector<int> &pf = perm[from->index];
          std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return pf[m->map[(size_t)m->dim * a]] < pf[m->map[(size_t)m->dim * b]];
          });
          pt.assign(to->size, 0);
          for (int k = 0; k < to->size; ++k) pt[order[k]] = k;
          break;
        }
    }
    if (pt.empty()) {  // a map from->to that is onto: first visit in new order
      for (op_map m : R.maps)
        if (m->to == to && m->from == from) {
          std::vector<int> cand(to->size, -1);
          int counter = 0;
          for (int i = 0; i < from->size; ++i)
            for (int d = 0; d < m->dim; ++d) {
              int idx = m->map[(size_t)iperm[from->index][i] * m->dim + d];
              if (cand[idx] == -1) cand[idx] = counter++;
            }
          if (counter == to->size) {
            pt.swap(cand);
            break;
          }
        }
    }
    if (pt.empty()) return;
    iperm[to->index].assign(to->size, 0);
    for (int i = 0; i < to->size; ++i) iperm[to->index][pt[i]] = i;
    for (op_map m : R.maps)
      if (m->to == to && perm[m->from->index].empty()) propagate(to, m->from);
    for (op_map m : R.maps)
      if (m->from == to && perm[m->to->index].empty()) propagate(to, m->to);
  }
};

}  // namespace

namespace {
void renumber_impl(op_map base) {
  Runtime &R = rt();
  op2h_flush_deferred();
  if (R.capturing) {
    printf("op_renumber error -- not allowed during graph capture\n");
    exit(-1);
  }
  if (!base || base->dim < 2) {
    printf("op_renumber error -- the base map must have dimension >= 2\n");
    exit(-1);
  }
  op2h_device_sync();
  for (op_dat d : R.dats)
    if (d && d->dirty_hd == 2) download_dat(d);
  // graph of base->to: two elements are adjacent when one element of
  // base->from references both
  op_set S = base->to;
  const int n = S->size;
  std::vector<std::vector<int>> nbr(n);
  for (int e = 0; e < base->from->size; ++e)
    for (int a = 0; a < base->dim; ++a)
      for (int b = 0; b < base->dim; ++b) {
        int u = base->map[(size_t)e * base->dim + a], v = base->map[(size_t)e * base->dim + b];
        if (u != v) nbr[u].push_back(v);
      }
  std::vector<int> off(n + 1, 0), adj;
  for (int v = 0; v < n; ++v) {
    std::sort(nbr[v].begin(), nbr[v].end());
    nbr[v].erase(std::unique(nbr[v].begin(), nbr[v].end()), nbr[v].end());
    off[v + 1] = off[v] + (int)nbr[v].size();
  }
  adj.reserve(off[n]);
  for (int v = 0; v < n; ++v) adj.insert(adj.end(), nbr[v].begin(), nbr[v].end());
  auto bandwidth = [&](double *avg) {
    const int *t = op2h::host_map(base);
    long long sum = 0;
    int mx = 0;
    for (int e = 0; e < base->from->size; ++e) {
      int dist = 0;
      for (int a = 0; a < base->dim; ++a)
        for (int b = 0; b < base->dim; ++b)
          dist = std::max(dist, std::abs(t[(size_t)e * base->dim + a] -
                                         t[(size_t)e * base->dim + b]));
      mx = std::max(mx, dist);
      sum += dist;
    }
    *avg = base->from->size ? (double)sum / base->from->size : 0.0;
    return mx;
  };
  double avg0, avg1;
  int max0 = bandwidth(&avg0);

  Renumber rn;
  rn.perm.assign(R.sets.size(), {});
  rn.iperm.assign(R.sets.size(), {});
  // the base ordering is built on the host unless OP_HIP_RCM=gpu; both
  // builds give the same permutation (the device build pays a host round trip
  // per BFS level, so the host one is faster at the sizes measured: 78 ms vs
  // 213 ms for 720k cells)
  const char *where = getenv("OP_HIP_RCM");
  const bool rcm_host = !where || strcmp(where, "gpu") != 0;
  if (where && strcmp(where, "host") && strcmp(where, "gpu")) {
    printf("op_renumber error -- OP_HIP_RCM must be gpu or host, got %s\n", where);
    exit(-1);
  }
  rn.perm[S->index] = rcm_host ? op2h::rcm_order(n, off, adj)
                               : op2h::rcm_order_gpu(n, off, adj, R.stream);
  rn.propagate(S, S);

  // reorder_set (op_renumber.cpp:153-198): rows of maps from the set and of
  // dats on it move; entries of maps into the set are renumbered.  Under
  // transparent renumbering the application's tables and rows stay as they
  // are: the maps' device tables are renumbered copies, and dats move to the
  // device numbering as they are uploaded.
  const bool keep = R.transparent;
  if (keep) {
    R.dev_maps.resize(R.maps.size());
    for (op_map m : R.maps)
      if (R.dev_maps[m->index].empty())
        R.dev_maps[m->index].assign(m->map, m->map + (size_t)m->from->size * m->dim);
  }
  auto table = [&](op_map m) { return keep ? R.dev_maps[m->index].data() : m->map; };
  for (op_set s : R.sets) {
    const std::vector<int> &p = rn.perm[s->index];
    if (p.empty() || s->size == 0) continue;
    for (op_map m : R.maps) {
      if (m->from == s) {
        int *t = table(m);
        std::vector<int> tmp(t, t + (size_t)s->size * m->dim);
        for (int i = 0; i < s->size; ++i)
          std::copy(tmp.begin() + (size_t)i * m->dim, tmp.begin() + (size_t)(i + 1) * m->dim,
                    t + (size_t)p[i] * m->dim);
      }
    }
    for (op_map m : R.maps)
      if (m->to == s) {
        int *t = table(m);
        for (size_t i = 0; i < (size_t)m->from->size * m->dim; ++i) t[i] = p[t[i]];
      }
    for (op_dat d : R.dats) {
      if (!d || d->set != s || keep) continue;
      std::vector<char> tmp(d->data, d->data + (size_t)d->size * s->size);
      for (int i = 0; i < s->size; ++i)
        std::memcpy(d->data + (size_t)p[i] * d->size, tmp.data() + (size_t)i * d->size,
                    d->size);
    }
    if (R.set_new_of_old.size() < R.sets.size()) R.set_new_of_old.resize(R.sets.size());
    std::vector<int> &acc = R.set_new_of_old[s->index];
    if (acc.empty()) {
      acc = p;
    } else {
      for (int &v : acc) v = p[v];
    }
  }
  for (op_map m : R.maps)
    HIPCHECK(hipMemcpy(m->map_d, table(m), (size_t)m->from->size * m->dim * sizeof(int),
                       hipMemcpyHostToDevice));
  for (op_dat d : R.dats)
    if (d) upload_dat(d);
  R.staged.clear();
  ++R.plan_epoch;
  R.color.clear();
  R.owner.clear();
  R.gather.clear();
  for (auto &kv : R.user_plans) (void)hipFree(kv.second.first);
  R.user_plans.clear();
  R.distinct_targets.clear();
  int max1 = bandwidth(&avg1);
  if (OP_diags > 0) {
    op_printf("Renumbering using base map %s (reverse Cuthill-McKee, built on the %s)\n",
              base->name, rcm_host ? "host" : "device");
    op_printf("Before renumbering: maximum bandwidth = %d average bandwidth = %.1f\n", max0, avg0);
    op_printf("After renumbering: maximum bandwidth = %d average bandwidth = %.1f\n", max1, avg1);
  }
}

// OP_HIP_RENUMBER: renumber every set reachable from the base map, once,
// before the first loop, and keep the inverse permutations for translating
// back to the application's numbering.  The base map is the one the token
// names, else the first declared map of dimension >= 2 from the largest set.
void transparent_renumber() {
  Runtime &R = rt();
  R.transparent_done = true;
  op_map base = nullptr;
  for (op_map m : R.maps) {
    if (!R.transparent_map.empty()) {
      if (R.transparent_map == m->name) base = m;
    } else if (m->dim >= 2 && (!base || m->from->size > base->from->size)) {
      base = m;
    }
  }
  if (!base) {
    if (!R.transparent_map.empty()) {
      printf("op2hip: OP_HIP_RENUMBER names map %s, which is not declared\n",
             R.transparent_map.c_str());
      exit(-1);
    }
    return;  // no map to renumber by
  }
  if (R.capturing) {
    printf("op2hip: OP_HIP_RENUMBER: the first loop must run before graph capture\n");
    exit(-1);
  }
  renumber_impl(base);
  R.set_old_of_new.assign(R.sets.size(), {});
  R.old_index_d.assign(R.sets.size(), nullptr);
  for (op_set st : R.sets) {
    if ((size_t)st->index >= R.set_new_of_old.size() || R.set_new_of_old[st->index].empty())
      continue;
    const std::vector<int> &p = R.set_new_of_old[st->index];
    std::vector<int> &q = R.set_old_of_new[st->index];
    q.assign(p.size(), 0);
    for (size_t i = 0; i < p.size(); ++i) q[p[i]] = (int)i;
    HIPCHECK(hipMalloc(&R.old_index_d[st->index], std::max<size_t>(q.size(), 1) * sizeof(int)));
    HIPCHECK(hipMemcpy(R.old_index_d[st->index], q.data(), q.size() * sizeof(int),
                       hipMemcpyHostToDevice));
  }
}

// Host data of a dat in the application's numbering: its own rows, which
// stay in that numbering under transparent renumbering too (uploads and
// downloads move rows between the numberings).
const char *app_rows(op_dat dat, std::vector<char> &) { return dat->data; }
}  // namespace

extern "C" void op_renumber(op_map base) {
  if (rt().transparent) {  // the device already works in its own numbering
    if (OP_diags > 0)
      op_printf("op_renumber: OP_HIP_RENUMBER is on, the call is a no-op\n");
    return;
  }
  renumber_impl(base);
}

extern "C" const int *op2h_set_old_index(op_set set) {
  Runtime &R = rt();
  if (!R.transparent || (size_t)set->index >= R.old_index_d.size()) return nullptr;
  return R.old_index_d[set->index];
}

namespace op2h {
const int *host_map(op_map m) {
  Runtime &R = rt();
  if ((size_t)m->index < R.dev_maps.size() && !R.dev_maps[m->index].empty())
    return R.dev_maps[m->index].data();
  return m->map;
}

const int *old_index_host(op_set set) {
  Runtime &R = rt();
  if (!R.transparent || (size_t)set->index >= R.set_old_of_new.size() ||
      R.set_old_of_new[set->index].empty())
    return nullptr;
  return R.set_old_of_new[set->index].data();
}
}  // namespace op2h

extern "C" int op2h_set_permutation(op_set set, int *new_of_old) {
  Runtime &R = rt();
  if (!R.transparent && (size_t)set->index < R.set_new_of_old.size() &&
      !R.set_new_of_old[set->index].empty()) {
    std::copy(R.set_new_of_old[set->index].begin(), R.set_new_of_old[set->index].end(),
              new_of_old);
    return 1;
  }
  for (int i = 0; i < set->size; ++i) new_of_old[i] = i;
  return 0;
}

// ===========================================================================
// backend controls
// ===========================================================================
extern "C" int op2h_set_loop_plan(const char *name, int plan, int block_size) {
  if (plan < 0 || plan >= OP2H_PLAN_COUNT) return -1;
  if (block_size < 0 || block_size > 1024 || block_size % 64) return -1;
  LoopConfig c;
  c.plan = plan;
  c.block_size = block_size;
  if (!name) {
    rt().all_cfg = c;
    rt().loop_cfg.clear();
  } else {
    rt().loop_cfg[name] = c;
  }
  return 0;
}

extern "C" int op2h_get_loop_plan(const char *name, int *plan, int *block_size) {
  LoopConfig c = rt().all_cfg;
  auto it = name ? rt().loop_cfg.find(name) : rt().loop_cfg.end();
  if (it != rt().loop_cfg.end()) c = it->second;
  if (plan) *plan = c.plan;
  if (block_size) *block_size = c.block_size;
  return 0;
}

extern "C" void op2h_set_sync_mode(int mode) { rt().sync_mode = mode; }
extern "C" int op2h_get_sync_mode(void) { return rt().sync_mode; }
extern "C" void op2h_set_plan_builder(int on_device) { rt().plan_on_device = on_device; }

extern "C" void op2h_enable_kernel_timing(int on) { rt().kernel_timing = on != 0; }

extern "C" void op2h_set_autotune(int on, int trials) {
  rt().autotune = on != 0;
  rt().tune_trials = std::max(1, trials);
}

extern "C" int op2h_tuning_done(void) {
  for (auto &kv : rt().tune)
    if (!kv.second.done) return 0;
  return 1;
}

extern "C" int op2h_tuning_records(int max, op2h_tuning_record *out) {
  int n = std::min((int)rt().tune_records.size(), std::max(max, 0));
  if (out)
    for (int i = 0; i < n; ++i) out[i] = rt().tune_records[i];
  return (int)rt().tune_records.size();
}

extern "C" int op2h_tuning_choice(const char *name, int *plan, int *block_size) {
  auto it = rt().tune.find(name);
  if (it == rt().tune.end() || !it->second.done) return -1;
  if (plan) *plan = it->second.plan;
  if (block_size) *block_size = it->second.block;
  return 0;
}

extern "C" int op2h_loop_stats(const char *name, int *calls, double *kernel_ms) {
  resolve_timers();
  for (int n = 0; n < OP_kern_max; ++n)
    if (OP_kernels[n].count && !std::strcmp(OP_kernels[n].name, name)) {
      auto it = rt().stats.find(n);
      if (calls) *calls = it == rt().stats.end() ? 0 : it->second.calls;
      if (kernel_ms) *kernel_ms = it == rt().stats.end() ? 0.0 : it->second.kernel_ms;
      return 0;
    }
  return -1;
}

extern "C" void op2h_reset_stats(void) {
  resolve_timers();
  rt().stats.clear();
  for (int n = 0; n < OP_kern_max; ++n) {
    OP_kernels[n].count = 0;
    OP_kernels[n].time = 0;
    OP_kernels[n].transfer = 0;
    OP_kernels[n].transfer2 = 0;
  }
}

extern "C" void *op2h_stream(void) {
  ensure_init();
  return rt().stream;
}

extern "C" void op2h_flush_deferred(void) {
  Runtime &R = rt();
  if (!R.pend_fn) return;
  op2h_flush_fn fn = R.pend_fn;
  void *arg = R.pend_arg;
  R.pend_fn = nullptr;
  R.pend_arg = nullptr;
  R.pend_tag.clear();
  fn(arg);
}

extern "C" void op2h_defer(const char *tag, op2h_flush_fn flush, void *arg) {
  Runtime &R = rt();
  op2h_flush_deferred();  // at most one pending launch
  R.pend_fn = flush;
  R.pend_arg = arg;
  R.pend_tag = tag ? tag : "";
}

extern "C" void op2h_accept_deferred(const char *tag) { rt().accept_tag = tag ? tag : ""; }

extern "C" int op2h_take_deferred(const char *tag, void **arg) {
  Runtime &R = rt();
  if (arg) *arg = nullptr;
  if (!R.pend_fn) return 0;
  if (!tag || R.pend_tag != tag) {
    op2h_flush_deferred();
    return 0;
  }
  if (arg) *arg = R.pend_arg;
  R.pend_fn = nullptr;
  R.pend_arg = nullptr;
  R.pend_tag.clear();
  ++R.fused;
  return 1;
}

extern "C" unsigned long long op2h_fused_count(void) { return rt().fused; }

extern "C" void op2h_set_fusion(int on) { rt().fusion = on != 0; }

// MI355X's last-level cache (the Infinity Cache, memory-side, shared by all
// XCDs): a stream larger than this cannot be served from it on the next pass.
static constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;

extern "C" void op2h_set_streaming(int mode) { rt().streaming = mode < 0 ? -1 : mode != 0; }

extern "C" int op2h_streaming(size_t stream_bytes) {
  const int m = rt().streaming;
  return m < 0 ? stream_bytes > kLastLevelCacheBytes : m;
}

extern "C" unsigned op2h_plan_epoch(void) { return rt().plan_epoch; }

extern "C" int op2h_fusion_enabled(void) {
  Runtime &R = rt();
  // fused launches only inside a captured graph, where no per-loop result,
  // time or tuning sample is taken between the loops
  return R.fusion && R.capturing && !R.kernel_timing && !R.autotune &&
         R.sync_mode != OP2H_SYNC_EVERY_LOOP && !op2h::profile_loops_on();
}

extern "C" void op2h_device_sync(void) {
  op2h_flush_deferred();
  if (!rt().initialised || rt().capturing) return;
  HIPCHECK(hipStreamSynchronize(rt().stream));
}

extern "C" const char *op2h_version(void) { return "op2hip 0.1 gfx950"; }

// ===========================================================================
// loop lifecycle
// ===========================================================================
namespace {
std::vector<op2h_const_hook> &const_hooks() {
  static std::vector<op2h_const_hook> hooks;  // filled by static initialisers
  return hooks;
}
}  // namespace

extern "C" void op2h_register_const_hook(op2h_const_hook hook) {
  if (hook) const_hooks().push_back(hook);
}

extern "C" void op_decl_const_char(int dim, char const *type, int size,
                                   char *dat, char const *name) {
  op2h_device_sync();
  int known = 0;  // every kernel file holding a copy of the constant gets it
  for (op2h_const_hook h : const_hooks()) known |= h(dim, type, size, dat, name);
  if (!known) {
    printf("op_decl_const error -- unknown constant %s\n", name);
    exit(-1);
  }
}

// ---- cost-model plan choice -------------------------------------------------
namespace {

// The loop as the cost model sees it (op2amd/costmodel.py describes the
// synthetic loops it is trained on the same way): the main map is the one of
// the first indirect written arg (else of the first indirect read one).
ModelLoop describe_loop(const char *name, op_set set, int nargs, const op_arg *args) {
  Runtime &R = rt();
  ModelLoop d;
  d.set_size = set->size;
  op_map main = nullptr;
  const op_arg *mainw = nullptr;
  bool inc = false, rw = false;
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || !a.opt || !a.map || a.idx < 0) continue;
    if (a.acc == OP_INC) inc = true;
    if (a.acc == OP_RW || a.acc == OP_WRITE) rw = true;
    if (a.acc != OP_READ && !mainw) mainw = &a;
  }
  if (mainw) {
    main = mainw->map;
  } else {
    for (int m = 0; m < nargs && !main; ++m)
      if (args[m].argtype == OP_ARG_DAT && args[m].opt && args[m].map && args[m].idx >= 0)
        main = args[m].map;
  }
  d.access = inc ? 1 : rw ? 2 : 0;
  // dim and precision: the written indirect dat, else the widest dat
  const op_arg *ref = mainw;
  for (int m = 0; m < nargs && !ref; ++m)
    if (args[m].argtype == OP_ARG_DAT && args[m].opt &&
        (!ref || args[m].dim > ref->dim))
      ref = &args[m];
  for (int m = 0; m < nargs; ++m)
    if (!mainw && args[m].argtype == OP_ARG_DAT && args[m].opt && args[m].dim > ref->dim)
      ref = &args[m];
  if (ref) {
    d.dim = std::max(1, ref->dim);
    d.is_double = type_code(ref->type) == 0 ? 1.0 : 0.0;
  }
  if (main) {
    d.target_size = main->to->size;
    std::vector<int> cols;
    for (int m = 0; m < nargs; ++m)
      if (args[m].argtype == OP_ARG_DAT && args[m].opt && args[m].map == main &&
          args[m].idx >= 0 && std::find(cols.begin(), cols.end(), args[m].idx) == cols.end())
        cols.push_back(args[m].idx);
    d.arity = std::min<int>(4, (int)cols.size());
    double sum = 0.0;
    for (int e = 0; e < set->size; ++e) {
      int mx = 0;
      for (int a : cols)
        for (int b : cols)
          mx = std::max(mx, std::abs(op2h::host_map(main)[(size_t)e * main->dim + a] -
                                     op2h::host_map(main)[(size_t)e * main->dim + b]));
      sum += mx;
    }
    d.bandwidth = set->size ? sum / set->size : 0.0;
  } else {
    d.target_size = set->size;
  }
  auto it = R.loop_arith.find(name);
  d.arith = it == R.loop_arith.end() ? 16 : it->second;
  return d;
}

// Feature vector of one variant, in the order of op2amd/costmodel.py's
// features(): sizes, shape, arithmetic, access, plan one-hot (direct, atomic,
// colour, staged, gather, owner), block (log and one-hot).
void model_features(const ModelLoop &d, int plan, int block, float *f) {
  int i = 0;
  f[i++] = (float)std::log10(std::max(1.0, d.set_size));
  f[i++] = (float)std::log10(std::max(1.0, d.target_size));
  f[i++] = (float)(d.arity / 4.0);
  f[i++] = (float)(std::log2(d.dim) / 3.0);
  f[i++] = (float)d.is_double;
  f[i++] = (float)(d.arith / 64.0);
  f[i++] = (float)(std::log10(1.0 + d.bandwidth) / 7.0);
  for (int a = 0; a < 3; ++a) f[i++] = d.access == a ? 1.f : 0.f;
  for (int p = 0; p < 6; ++p) f[i++] = plan == p ? 1.f : 0.f;
  f[i++] = (float)(std::log2((double)block) / 10.0);
  for (int b : {64, 128, 256, 512, 1024}) f[i++] = block == b ? 1.f : 0.f;
}

// The loop's variants, best predicted first (plan 0 = direct: the loop's
// own plan, block size only).
std::vector<std::pair<int, int>> model_rank(const char *name, op_set set, int nargs,
                                            const op_arg *args, unsigned plan_mask) {
  Runtime &R = rt();
  const ModelLoop d = describe_loop(name, set, nargs, args);
  R.model_loops[name] = d;
  std::vector<std::pair<int, int>> cands;
  const int all_blocks[5] = {64, 128, 256, 512, 1024};
  if (d.access == 0) {
    for (int b : all_blocks) cands.push_back({0, b});
  } else {
    for (int p : {OP2H_PLAN_ATOMIC, OP2H_PLAN_COLOR, OP2H_PLAN_STAGED, OP2H_PLAN_GATHER,
                  OP2H_PLAN_OWNER}) {
      if (!((plan_mask >> p) & 1)) continue;
      if (p == OP2H_PLAN_ATOMIC && d.access != 1) continue;  // increments only
      if (p == OP2H_PLAN_OWNER && d.access != 1) continue;
      for (int b : all_blocks) {
        if (p == OP2H_PLAN_STAGED && (b < 128 || b > 512)) continue;
        if (p == OP2H_PLAN_OWNER && b > 256) continue;  // LDS of larger patches
        cands.push_back({p, b});
      }
    }
  }
  if (cands.empty()) return cands;
  const int nf = OP2H_MODEL_FEATURES, n = (int)cands.size();
  std::vector<float> X((size_t)n * nf), y(n), acc(n, 0.f);
  for (int i = 0; i < n; ++i) {
    float *f = &X[(size_t)i * nf];
    model_features(d, cands[i].first, cands[i].second, f);
    for (int k = 0; k < nf; ++k) f[k] = (f[k] - R.model.mean[k]) / R.model.scale[k];
  }
  for (op2amd_mlp *m : R.model.members) {
    if (op2amd_mlp_forward(m, X.data(), n, y.data())) {
      printf("op2hip: cost model evaluation failed for %s\n", name);
      exit(-1);
    }
    for (int i = 0; i < n; ++i) acc[i] += y[i];
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return acc[a] < acc[b]; });
  std::vector<std::pair<int, int>> ranked;
  for (int i : order) ranked.push_back(cands[i]);
  return ranked;
}

}  // namespace

extern "C" int op2h_set_plan_model(struct op2amd_mlp *const *members, int n_members,
                                   int nfeat, const float *mean, const float *scale,
                                   int top) {
  Runtime &R = rt();
  if (n_members <= 0) {
    R.model = PlanModel();
    return 0;
  }
  if (!members || nfeat != OP2H_MODEL_FEATURES || !mean || !scale || top < 1) return -1;
  for (int k = 0; k < nfeat; ++k)
    if (!(scale[k] > 0.f)) return -1;
  R.model.members.assign(members, members + n_members);
  R.model.mean.assign(mean, mean + nfeat);
  R.model.scale.assign(scale, scale + nfeat);
  R.model.top = top;
  return 0;
}

extern "C" void op2h_set_loop_arith(const char *name, int per_element) {
  rt().loop_arith[name] = std::min(64, std::max(4, per_element));
}

extern "C" int op2h_plan_model_loop(const char *name, double *desc8) {
  auto it = rt().model_loops.find(name);
  if (it == rt().model_loops.end() || !desc8) return -1;
  const ModelLoop &d = it->second;
  const double v[8] = {d.set_size, d.target_size, d.arity, d.dim,
                       d.is_double, d.arith, d.bandwidth, (double)d.access};
  std::copy(v, v + 8, desc8);
  return 0;
}

extern "C" void op2h_loop_begin(op2h_loop_ctx *ctx, int kernel_idx,
                                const char *name, op_set set, int nargs,
                                op_arg *args, int def_plan, int def_block) {
  unsigned all = 0;
  for (int p = 1; p < OP2H_PLAN_COUNT; ++p) all |= 1u << p;
  op2h_loop_begin_mask(ctx, kernel_idx, name, set, nargs, args, def_plan,
                       def_block,
                       def_plan == OP2H_PLAN_STAGED ? all : 1u << def_plan);
}

namespace {
// Plan and block size of one call: a recorded per-loop or all-loop setting,
// else the autotuner's or the cost model's choice, else the stub's default.
void choose_plan(Runtime &R, op2h_loop_ctx *ctx, const char *name, op_set set,
                 int nargs, op_arg *args, int def_plan, int def_block,
                 unsigned plan_mask) {
  LoopConfig c = R.all_cfg;
  auto it = R.loop_cfg.find(name);
  if (it != R.loop_cfg.end()) c = it->second;
  const bool with_model = !R.model.members.empty();
  if ((R.autotune || with_model) && c.plan == OP2H_PLAN_AUTO && c.block_size == 0 &&
      set->size > 0) {
    TuneState &ts = R.tune[name];
    if (ts.cands.empty() && with_model && !R.capturing) {
      // the cost model ranks the variants: the best one, or the `top` best to
      // be timed by the autotuner's trials
      std::vector<std::pair<int, int>> ranked = model_rank(name, set, nargs, args, plan_mask);
      for (auto &v : ranked)
        if (v.first == 0) v.first = def_plan;  // conflict-free: the loop's own plan
      if (!ranked.empty()) {
        const size_t k = std::min(ranked.size(), (size_t)R.model.top);
        ts.cands.assign(ranked.begin(), ranked.begin() + k);
        ts.best_ms.assign(ts.cands.size(), 1e300);
        if (k == 1) {
          ts.done = true;
          ts.plan = ts.cands[0].first;
          ts.block = ts.cands[0].second;
        }
      }
    }
    if (ts.cands.empty() && R.capturing) {  // first seen inside a capture
      ctx->plan = c.plan == OP2H_PLAN_AUTO ? def_plan : c.plan;
      ctx->block_size = def_block;
      ctx->part_size = ctx->block_size;
      return;
    }
    if (ts.cands.empty()) {
      int nplans = 0;
      for (int p = 1; p < OP2H_PLAN_COUNT; ++p) nplans += (plan_mask >> p) & 1;
      if (nplans > 1) {
        for (int p = 1; p < OP2H_PLAN_COUNT; ++p)
          if ((plan_mask >> p) & 1)
            for (int b : {128, 256, 512}) ts.cands.push_back({p, b});
      } else {
        for (int b : {128, 256, 512, 1024}) ts.cands.push_back({def_plan, b});
      }
      ts.best_ms.assign(ts.cands.size(), 1e300);
    }
    if (!ts.done && !R.capturing) {
      ctx->plan = ts.cands[ts.next].first;
      ctx->block_size = ts.cands[ts.next].second;
      ctx->part_size = ctx->block_size;
      ctx->tuning = 1;
      return;
    }
    if (!ts.done) {  // capturing before tuning finished: best so far
      size_t bi = 0;
      for (size_t i = 0; i < ts.best_ms.size(); ++i)
        if (ts.best_ms[i] < ts.best_ms[bi]) bi = i;
      ctx->plan = ts.cands[bi].first;
      ctx->block_size = ts.cands[bi].second;
      ctx->part_size = ctx->block_size;
      return;
    }
    ctx->plan = ts.plan;
    ctx->block_size = ts.block;
    ctx->part_size = ctx->block_size;
    return;
  }
  ctx->plan = c.plan == OP2H_PLAN_AUTO ? def_plan : c.plan;
  if (!((plan_mask >> ctx->plan) & 1)) {
    if (it != R.loop_cfg.end()) {
      printf("op2hip: loop %s does not implement plan %d\n", name, ctx->plan);
      exit(-1);
    }
    ctx->plan = def_plan;  // a setting for all loops: this loop keeps its own
  }
  ctx->block_size = c.block_size ? c.block_size
                                 : (OP_block_size > 0 && OP_block_size % 64 == 0 &&
                                            OP_block_size != 256
                                        ? OP_block_size
                                        : def_block);
  ctx->part_size = OP_part_size > 0 ? OP_part_size : ctx->block_size;
}
}  // namespace

extern "C" void op2h_loop_begin_mask(op2h_loop_ctx *ctx, int kernel_idx,
                                     const char *name, op_set set, int nargs,
                                     op_arg *args, int def_plan, int def_block,
                                     unsigned plan_mask) {
  Runtime &R = rt();
  ensure_init();
  if (R.transparent && !R.transparent_done) transparent_renumber();
  if (R.pend_fn && R.accept_tag != R.pend_tag) op2h_flush_deferred();
  R.accept_tag.clear();
  std::memset(ctx, 0, sizeof(*ctx));
  roctxRangePushA(name);
  ctx->kernel_idx = kernel_idx;
  ctx->name = name;
  ctx->set = set;
  ctx->nargs = nargs;
  ctx->args = args;
  ctx->stream = R.stream;
  ctx->capturing = R.capturing ? 1 : 0;
  op_timing_realloc(kernel_idx);
  OP_kernels[kernel_idx].name = name;
  OP_kernels[kernel_idx].count += 1;
  R.prof_loop = !R.capturing && op2h::profile_loops_on();
  if (R.prof_loop) op2h_device_sync();
  ctx->wall0 = wall_now();
  if (OP_diags > 0) {
    int ninds = 0;
    for (int m = 0; m < nargs; ++m) op_arg_check(set, m, args[m], &ninds, name);
  }
  if (OP_diags > 2) printf(" kernel routine: %s\n", name);
  for (int m = 0; m < nargs; ++m) {
    op_arg &a = args[m];
    if (a.argtype == OP_ARG_DAT && a.opt && a.dat && a.dat->dirty_hd == 1) {
      if (R.capturing) {
        printf("op2hip: dat %s changed on the host during graph capture\n",
               a.dat->name);
        exit(-1);
      }
      upload_dat(a.dat);
    }
    if (a.argtype == OP_ARG_DAT && a.dat) a.data_d = a.dat->data_d;
  }
  choose_plan(R, ctx, name, set, nargs, args, def_plan, def_block, plan_mask);
  if (R.prof_loop) {
    if (op2h::profile_level() >= 3) {
      resolve_timers();
      R.prof_kms0 = R.stats[kernel_idx].kernel_ms;
    }
    R.prof_comp_t0 = wall_now();
  }
}

extern "C" void op2h_kernel_timer_start(op2h_loop_ctx *ctx) {
  Runtime &R = rt();
  if (ctx->tuning) {
    if (!R.tune_ev[0]) {
      HIPCHECK(hipEventCreate(&R.tune_ev[0]));
      HIPCHECK(hipEventCreate(&R.tune_ev[1]));
    }
    HIPCHECK(hipEventRecord(R.tune_ev[0], R.stream));
    return;
  }
  if (!R.timing_on() || R.capturing) return;
  PendingTimer t{ctx->kernel_idx, take_event(), take_event()};
  HIPCHECK(hipEventRecord(t.a, R.stream));
  R.pending.push_back(t);
}

extern "C" int op2h_kernel_timer_events(op2h_loop_ctx *ctx, void **start,
                                        void **stop) {
  Runtime &R = rt();
  *start = *stop = nullptr;
  if (ctx->tuning || !R.timing_on() || R.capturing) return 0;
  PendingTimer t{ctx->kernel_idx, take_event(), take_event()};
  R.pending.push_back(t);
  *start = t.a;
  *stop = t.b;
  return 1;
}

extern "C" void op2h_kernel_timer_stop(op2h_loop_ctx *ctx) {
  Runtime &R = rt();
  if (ctx->tuning) {
    HIPCHECK(hipEventRecord(R.tune_ev[1], R.stream));
    HIPCHECK(hipEventSynchronize(R.tune_ev[1]));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, R.tune_ev[0], R.tune_ev[1]));
    TuneState &ts = R.tune[ctx->name];
    op2h_tuning_record rec;
    std::memset(&rec, 0, sizeof(rec));
    std::snprintf(rec.loop, sizeof rec.loop, "%s", ctx->name);
    rec.set_size = ctx->set->size;
    rec.plan = ctx->plan;
    rec.block_size = ctx->block_size;
    rec.kernel_ms = ms;
    // the first call of a variant also builds its plan: it is not timed
    if (ts.trials_done > 0) {
      R.tune_records.push_back(rec);
      ts.best_ms[ts.next] = std::min(ts.best_ms[ts.next], (double)ms);
    }
    if (++ts.trials_done > R.tune_trials) {
      ts.trials_done = 0;
      if (++ts.next == ts.cands.size()) {
        size_t bi = 0;
        for (size_t i = 0; i < ts.best_ms.size(); ++i)
          if (ts.best_ms[i] < ts.best_ms[bi]) bi = i;
        ts.done = true;
        ts.plan = ts.cands[bi].first;
        ts.block = ts.cands[bi].second;
        if (OP_diags > 1)
          printf(" autotune %s: plan %d block %d (%.2f us)\n", ctx->name, ts.plan,
                 ts.block, ts.best_ms[bi] * 1e3);
      }
    }
    return;
  }
  if (!R.timing_on() || R.capturing || R.pending.empty()) return;
  (void)ctx;
  HIPCHECK(hipEventRecord(R.pending.back().b, R.stream));
}

extern "C" void op2h_loop_end(op2h_loop_ctx *ctx) {
  Runtime &R = rt();
  for (int m = 0; m < ctx->nargs; ++m) {
    op_arg &a = ctx->args[m];
    if (a.argtype == OP_ARG_DAT && a.opt && a.acc != OP_READ) a.dat->dirty_hd = 2;
  }
  // gbl results
  bool need_sync = R.sync_mode == OP2H_SYNC_EVERY_LOOP;
  for (int g = 0; g < ctx->ngbl; ++g) {
    op_arg &a = ctx->args[ctx->gbl_arg[g]];
    int info = -1;
    for (int m = 0; m < ctx->nargs; ++m)  // the info arg names its gbl in acc
      if (ctx->args[m].argtype == OP_ARG_INFO && ctx->args[m].opt &&
          ctx->args[m].acc == ctx->gbl_arg[g])
        info = m;
    bool with_info = (a.acc == OP_MIN || a.acc == OP_MAX) && info >= 0;
    size_t bytes = (size_t)a.size + (with_info ? ctx->args[info].size : 0);
    // inside a captured graph the slots reach the host through the history
    // record of the launch (no per-loop copy node)
    if (!R.capturing)
      HIPCHECK(hipMemcpyAsync(ctx->gbl_slot_h[g], ctx->gbl_slot_d[g], bytes,
                              hipMemcpyDeviceToHost, R.stream));
    if (R.capturing) {
      DeferredGbl d;
      d.acc = a.acc;
      d.type = type_code(a.type);
      d.dim = a.dim;
      d.host_ptr = a.data;
      d.snapshot.assign(a.data, a.data + a.size);
      d.info_ptr = with_info ? ctx->args[info].data : nullptr;
      d.info_size = with_info ? ctx->args[info].size : 0;
      if (with_info)
        d.info_snapshot.assign(ctx->args[info].data,
                               ctx->args[info].data + d.info_size);
      d.slot_h = (char *)ctx->gbl_slot_h[g];
      d.bytes = bytes;
      R.deferred.push_back(std::move(d));
      R.history_sources.push_back({ctx->gbl_slot_d[g], bytes});
    } else {
      need_sync = true;
    }
  }
  if (!R.capturing) {
    if (need_sync) HIPCHECK(hipStreamSynchronize(R.stream));
    for (int g = 0; g < ctx->ngbl; ++g) {
      op_arg &a = ctx->args[ctx->gbl_arg[g]];
      int info = -1;
      for (int m = 0; m < ctx->nargs; ++m)
        if (ctx->args[m].argtype == OP_ARG_INFO && ctx->args[m].opt &&
            ctx->args[m].acc == ctx->gbl_arg[g])
          info = m;
      bool with_info = (a.acc == OP_MIN || a.acc == OP_MAX) && info >= 0;
      const char *res = (const char *)ctx->gbl_slot_h[g];
      combine_gbl(a.acc, type_code(a.type), a.dim, a.data, res,
                  with_info ? ctx->args[info].data : nullptr,
                  with_info ? res + a.size : nullptr,
                  with_info ? ctx->args[info].size : 0);
    }
  }
  if (R.prof_loop) {
    op2h_device_sync();
    double kms = -1.0;
    if (op2h::profile_level() >= 3) {
      resolve_timers();
      kms = R.stats[ctx->kernel_idx].kernel_ms - R.prof_kms0;
    }
    const double t = wall_now();
    static const char *plan_name[OP2H_PLAN_COUNT] = {"auto",   "atomic", "color",
                                                     "staged", "gather", "owner"};
    bool indirect = false;
    for (int m = 0; m < ctx->nargs; ++m)
      indirect |= ctx->args[m].argtype == OP_ARG_DAT && ctx->args[m].opt &&
                  ctx->args[m].map != nullptr;
    const char *variant = !indirect ? "direct"
                          : ctx->plan > 0 && ctx->plan < OP2H_PLAN_COUNT ? plan_name[ctx->plan]
                                                                          : "auto";
    op2h::profile_loop(ctx->name, variant, (long long)((t - ctx->wall0) * 1e9),
                       (long long)((t - R.prof_comp_t0) * 1e9), kms);
    R.prof_loop = false;
  }
  double t1 = wall_now();
  OP_kernels[ctx->kernel_idx].time += t1 - ctx->wall0;
  OP_kernels[ctx->kernel_idx].times[0] += t1 - ctx->wall0;
  roctxRangePop();
}

extern "C" int op2h_dat_soa_stride(op_dat dat) {
  return dat_is_soa(dat) ? dat->set->stride : 0;
}

extern "C" void *op2h_gbl_slot(op2h_loop_ctx *ctx, int argi, int info_argi) {
  Runtime &R = rt();
  if (ctx->ngbl >= OP2H_MAX_GBL) {
    printf("op2hip: too many gbl reductions in loop %s\n", ctx->name);
    exit(-1);
  }
  size_t bytes = (size_t)ctx->args[argi].size +
                 (info_argi >= 0 ? (size_t)ctx->args[info_argi].size : 0);
  bytes = (bytes + 63) & ~(size_t)63;
  std::pair<void *, void *> slot;
  if (bytes > 256) {
    printf("op2hip: gbl argument too large (%zu bytes)\n", bytes);
    exit(-1);
  }
  if (R.capturing) {
    if (R.capture_slot_next >= (int)R.capture_slots.size()) {
      printf("op2hip: too many gbl results in one captured graph\n");
      exit(-1);
    }
    slot = R.capture_slots[R.capture_slot_next++];
  } else {
    auto key = std::make_pair(ctx->kernel_idx, argi);
    auto it = R.slots.find(key);
    if (it == R.slots.end()) {
      HIPCHECK(hipMalloc(&slot.first, 256));
      HIPCHECK(hipHostMalloc(&slot.second, 256, hipHostMallocDefault));
      R.slots[key] = slot;
    } else {
      slot = it->second;
    }
  }
  int g = ctx->ngbl++;
  ctx->gbl_arg[g] = argi;
  ctx->gbl_slot_d[g] = slot.first;
  ctx->gbl_slot_h[g] = slot.second;
  return slot.first;
}

extern "C" int op2h_gbl_partials(op2h_loop_ctx *ctx, int argi, int info_argi,
                                 int nparts, double **vals, int **locs) {
  Runtime &R = rt();
  *vals = nullptr;
  *locs = nullptr;
  const op_arg &a = ctx->args[argi];
  if (!R.capturing || !R.partial_pool || nparts < 1 ||
      nparts > op2h_red::kRedThreads * op2h_red::kRedMaxLoads || a.dim != 1 ||
      type_code(a.type) != 0)
    return 0;
  const bool extremum = a.acc == OP_MIN || a.acc == OP_MAX;
  if (a.acc != OP_INC && !extremum) return 0;
  if (extremum && info_argi >= 0 &&
      (ctx->args[info_argi].dim != 1 || type_code(ctx->args[info_argi].type) != 2))
    return 0;
  const size_t need = (size_t)nparts * (sizeof(double) + (extremum ? sizeof(int) : 0));
  const size_t at = (R.partial_pool_next + 255) / 256 * 256;
  if (at + need > R.partial_pool_bytes) {
    printf("op2hip: deferred reductions of one graph exceed %zu bytes\n",
           R.partial_pool_bytes);
    exit(-1);
  }
  R.partial_pool_next = at + need;
  void *slot = op2h_gbl_slot(ctx, argi, info_argi);
  Runtime::Partials P;
  P.kind = a.acc == OP_INC ? 1 : a.acc == OP_MAX ? 2 : 3;
  P.nparts = nparts;
  P.vals = (const double *)(R.partial_pool + at);
  P.locs = extremum ? (const int *)(R.partial_pool + at + (size_t)nparts * sizeof(double))
                    : nullptr;
  R.partial_of_slot[slot] = P;
  *vals = (double *)P.vals;
  *locs = (int *)P.locs;
  return 1;
}

extern "C" void *op2h_reduct_scratch(size_t bytes) {
  Runtime &R = rt();
  if (bytes > R.red_scratch_bytes) {
    if (R.capturing) {
      printf("op2hip: reduction scratch must be sized before graph capture\n");
      exit(-1);
    }
    if (R.red_scratch) {
      HIPCHECK(hipStreamSynchronize(R.stream));
      HIPCHECK(hipFree(R.red_scratch));
    }
    size_t nb = std::max(bytes, (size_t)1 << 16);
    HIPCHECK(hipMalloc(&R.red_scratch, nb));
    R.red_scratch_bytes = nb;
  }
  return R.red_scratch;
}

extern "C" unsigned int *op2h_reduct_counter(void) { return rt().red_counter; }

extern "C" double op2h_algorithmic_bytes(op_set set, int nargs,
                                         const op_arg *args) {
  std::map<op_map, long long> &distinct = rt().distinct_targets;
  double bytes = 0.0;
  std::vector<op_dat> seen;
  std::vector<std::pair<op_map, int>> seen_maps;
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || !a.opt || !a.dat) continue;
    if (a.map && a.idx >= 0 &&
        std::find(seen_maps.begin(), seen_maps.end(), std::make_pair(a.map, a.idx)) ==
            seen_maps.end()) {
      seen_maps.push_back({a.map, a.idx});
      bytes += 4.0 * a.map->from->size;
    }
    if (std::find(seen.begin(), seen.end(), a.dat) != seen.end()) continue;
    seen.push_back(a.dat);
    double factor = (a.acc == OP_INC || a.acc == OP_RW) ? 2.0 : 1.0;
    long long count = set->size;
    if (a.map) {
      auto it = distinct.find(a.map);
      if (it == distinct.end()) {
        std::vector<char> hit(a.map->to->size, 0);
        long long c = 0;
        for (size_t i = 0; i < (size_t)a.map->from->size * a.map->dim; ++i)
          if (!hit[a.map->map[i]]) {
            hit[a.map->map[i]] = 1;
            ++c;
          }
        it = distinct.emplace(a.map, c).first;
      }
      count = it->second;
    }
    bytes += factor * (double)count * a.dat->size;
  }
  return bytes;
}

// ===========================================================================
// plans
// ===========================================================================
namespace {
// A plan depends on the loop's set, maps, indices and access modes, on which
// arguments name the same dat, and on the dats' shapes, not on the dats
// themselves: the key names each dat by the first argument that passes it,
// so a loop over temporary dats declared afresh every iteration
// (op_decl_dat_temp) finds the plan its first iteration built.
std::string arg_signature(const char *name, op_set set, int nargs,
                          const op_arg *args, int part) {
  std::string s = std::string(name) + "|" + std::to_string(set->index) + "|" +
                  std::to_string(part);
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    int first = -1;
    if (a.dat)
      for (int k = 0; k <= m && first < 0; ++k)
        if (args[k].dat == a.dat) first = k;
    s += "|" + std::to_string(a.argtype) + "," + std::to_string(first) + "," +
         (a.dat ? std::to_string(a.dat->set->index) + ":" + std::to_string(a.dat->dim) +
                      ":" + std::to_string(a.dat->size) + ":" + std::to_string(dat_is_soa(a.dat))
                : std::string("-")) +
         "," + std::to_string(a.map ? a.map->index : -1) + "," + std::to_string(a.idx) +
         "," + std::to_string(a.acc) + "," + std::to_string(a.opt);
  }
  return s;
}
}  // namespace

extern "C" const op2h_staged_plan *op2h_plan_staged_for(const char *name,
                                                        op_set set,
                                                        int part_size, int nargs,
                                                        op_arg *args) {
  Runtime &R = rt();
  std::string sig = arg_signature(name, set, nargs, args, part_size);
  for (auto &p : R.staged)
    if (p->signature == sig) return &p->plan;
  if (R.capturing) {
    printf("op2hip: plan for %s must be built before graph capture\n", name);
    exit(-1);
  }
  double t0 = wall_now();
  auto rec = op2h::build_staged_plan(set, part_size, nargs, args,
                                     R.plan_on_device, R.stream);
  rec->signature = sig;
  rec->plan.build_ms = (wall_now() - t0) * 1e3;
  if (OP_diags > 0) op2h::check_staged_or_exit(*rec, name, set, nargs, args);
  R.staged.push_back(std::move(rec));
  return &R.staged.back()->plan;
}

extern "C" const op2h_staged_plan *op2h_plan_staged(op2h_loop_ctx *ctx) {
  const op2h_staged_plan *p = op2h_plan_staged_for(
      ctx->name, ctx->set, ctx->part_size, ctx->nargs, ctx->args);
  OP_kernels[ctx->kernel_idx].plan_time += 0.0f;
  return p;
}

extern "C" const op2h_color_plan *op2h_plan_color_for(const char *name,
                                                      op_set set, int nargs,
                                                      op_arg *args) {
  Runtime &R = rt();
  std::string sig = arg_signature(name, set, nargs, args, -1);
  for (auto &p : R.color)
    if (p->signature == sig) return &p->plan;
  if (R.capturing) {
    printf("op2hip: plan for %s must be built before graph capture\n", name);
    exit(-1);
  }
  double t0 = wall_now();
  auto rec = op2h::build_color_plan(set, nargs, args, R.plan_on_device, R.stream);
  rec->signature = sig;
  rec->plan.build_ms = (wall_now() - t0) * 1e3;
  if (OP_diags > 0) op2h::check_color_or_exit(*rec, name, set, nargs, args);
  R.color.push_back(std::move(rec));
  return &R.color.back()->plan;
}

extern "C" const op2h_color_plan *op2h_plan_color(op2h_loop_ctx *ctx) {
  return op2h_plan_color_for(ctx->name, ctx->set, ctx->nargs, ctx->args);
}

extern "C" const op2h_gather_plan *op2h_plan_gather(op_map map) {
  Runtime &R = rt();
  auto it = R.gather.find(map);
  if (it != R.gather.end()) return &it->second->plan;
  if (R.capturing) {
    printf("op2hip: plan for map %s must be built before graph capture\n",
           map->name);
    exit(-1);
  }
  double t0 = wall_now();
  auto rec = op2h::build_gather_plan(map, R.stream);
  rec->plan.build_ms = (wall_now() - t0) * 1e3;
  if (OP_diags > 0) op2h::check_gather_or_exit(*rec, map);
  auto *p = &rec->plan;
  R.gather[map] = std::move(rec);
  return p;
}

namespace {
const op2h_owner_plan *owner_plan_for(const char *name, op_set set, int block_cells,
                                      int nargs, op_arg *args, bool key_instances) {
  Runtime &R = rt();
  // device-ordered instances make a plan of their own only where the two
  // orders differ (transparent renumbering)
  if (!op2h::old_index_host(set)) key_instances = true;
  std::string sig = arg_signature(name, set, nargs, args, block_cells) +
                    (key_instances ? "" : "/local");
  for (auto &p : R.owner)
    if (p->signature == sig) return &p->plan;
  if (R.capturing) {
    printf("op2hip: plan for %s must be built before graph capture\n", name);
    exit(-1);
  }
  double t0 = wall_now();
  std::string err;
  auto rec = op2h::build_owner_plan(set, block_cells, nargs, args, R.stream, err,
                                    key_instances);
  if (!rec) {
    printf("op2hip: %s: %s\n", name, err.c_str());
    return nullptr;
  }
  rec->signature = sig;
  rec->plan.build_ms = (wall_now() - t0) * 1e3;
  if (OP_diags > 0) op2h::check_owner_or_exit(*rec, name, nargs, args);
  R.owner.push_back(std::move(rec));
  return &R.owner.back()->plan;
}
}  // namespace

extern "C" const op2h_owner_plan *op2h_plan_owner_for(const char *name, op_set set,
                                                      int block_cells, int nargs,
                                                      op_arg *args) {
  return owner_plan_for(name, set, block_cells, nargs, args, true);
}

extern "C" const op2h_owner_plan *op2h_plan_owner(op2h_loop_ctx *ctx) {
  return op2h_plan_owner_for(ctx->name, ctx->set, ctx->part_size, ctx->nargs,
                             ctx->args);
}

extern "C" const op2h_owner_plan *op2h_plan_owner_local(op2h_loop_ctx *ctx) {
  return owner_plan_for(ctx->name, ctx->set, ctx->part_size, ctx->nargs, ctx->args, false);
}

extern "C" int op2h_owner_plan_download(const op2h_owner_plan *p,
                                        int *owned_offs, int *owned,
                                        int *elem_offs, int *elems,
                                        int *csr_offs, unsigned short *csr) {
  for (auto &r : rt().owner)
    if (&r->plan == p)
      return op2h::download_owner(*r, owned_offs, owned, elem_offs, elems,
                                  csr_offs, csr);
  return 1;
}

extern "C" void *op2h_plan_user(const char *key, size_t bytes, int *fresh) {
  Runtime &R = rt();
  auto it = R.user_plans.find(key);
  if (it != R.user_plans.end() && it->second.second == bytes) {
    *fresh = 0;
    return it->second.first;
  }
  if (R.capturing) {
    printf("op2hip: plan %s must be built before graph capture\n", key);
    exit(-1);
  }
  if (it != R.user_plans.end()) {
    HIPCHECK(hipStreamSynchronize(R.stream));
    (void)hipFree(it->second.first);
    R.user_plans.erase(it);
  }
  void *p = nullptr;
  HIPCHECK(hipMalloc(&p, std::max(bytes, (size_t)16)));
  R.user_plans[key] = {p, bytes};
  *fresh = 1;
  return p;
}

extern "C" int op2h_staged_plan_download(const op2h_staged_plan *p, int *thrcol,
                                         int *blkcol, int *offset, int *nelems) {
  return op2h::download_staged(p, thrcol, blkcol, offset, nelems);
}

extern "C" int op2h_color_plan_download(const op2h_color_plan *p,
                                        int *color_of_elem) {
  return op2h::download_color(p, color_of_elem);
}

// ===========================================================================
// graph capture
// ===========================================================================
namespace {
constexpr int kSlotPool = 128;
constexpr int kHistSources = 64;
constexpr size_t kHistoryBytes = 64u << 20;

struct HistSources {
  const unsigned long long *p[kHistSources];  // slot words, or partial values
  const int *locs[kHistSources];              // partial payloads (max/min)
  int words[kHistSources];
  int off[kHistSources];     // first word of the source in a record
  int kind[kHistSources];    // 0 copy the slot, 1 sum, 2 max, 3 min of partials
  int nparts[kHistSources];
  int n;
  int red[kHistSources];     // indices of the deferred reductions
  int nred;
};

// Appends the gbl results recorded in a graph to the device history: record
// r holds all result words of launch r, in record order.  One workgroup of
// op2h_red::kRedThreads per deferred reduction (at least one): workgroup b
// finishes reduction b from its partials with op2h_red::group_reduce (the
// eager finish kernels' reduction, so results are the same bits), and
// workgroup 0 also copies the slot sources.  Every workgroup reads the launch
// counter first; the last one to finish (a ticket on `done`, after a release
// fence that orders its reads and writes before it) advances the counter and
// resets the ticket for the next launch.
__global__ void __launch_bounds__(op2h_red::kRedThreads)
    history_append(HistSources src, double *hist, unsigned long long *count,
                   unsigned int *done, int capacity, int total_words) {
  __shared__ double s_v[op2h_red::kRedThreads / 64];
  __shared__ int s_l[op2h_red::kRedThreads / 64];
  const unsigned long long c = *count;
  const size_t base = (size_t)(c % (unsigned long long)capacity) * total_words;
  const int t = threadIdx.x, b = blockIdx.x;
  if (b == 0)
    for (int s = 0; s < src.n; ++s)
      if (src.kind[s] == 0)
        for (int k = t; k < src.words[s]; k += blockDim.x)
          hist[base + src.off[s] + k] = __longlong_as_double((long long)src.p[s][k]);
  if (b < src.nred) {
    const int s = src.red[b];
    const op2h_red::ValLoc v = op2h_red::group_reduce(
        reinterpret_cast<const double *>(src.p[s]), src.locs[s], src.nparts[s], src.kind[s], t,
        s_v, s_l);
    if (t == 0) {
      hist[base + src.off[s]] = v.v;
      if (src.words[s] > 1)  // the info int in the low half of the next word
        hist[base + src.off[s] + 1] = __longlong_as_double((long long)(unsigned int)v.loc);
    }
  }
  __syncthreads();
  if (t == 0) {
    __threadfence();
    if (atomicAdd(done, 1u) == gridDim.x - 1) {
      *done = 0u;
      *count = c + 1;
    }
  }
}
}  // namespace

extern "C" int op2h_graph_begin(void) {
  Runtime &R = rt();
  ensure_init();
  if (R.capturing) return -1;
  op2h_flush_deferred();
  op2h_graph_destroy();
  resolve_timers();
  HIPCHECK(hipStreamSynchronize(R.stream));
  R.deferred.clear();
  R.history_sources.clear();
  // everything a capture needs is allocated before capture starts
  for (int i = 0; i < kSlotPool; ++i) {
    std::pair<void *, void *> slot;
    HIPCHECK(hipMalloc(&slot.first, 256));
    HIPCHECK(hipHostMalloc(&slot.second, 256, hipHostMallocDefault));
    R.capture_slots.push_back(slot);
  }
  R.capture_slot_next = 0;
  R.partial_pool_bytes = 16u << 20;
  HIPCHECK(hipMalloc(&R.partial_pool, R.partial_pool_bytes));
  R.partial_pool_next = 0;
  R.partial_of_slot.clear();
  HIPCHECK(hipMalloc(&R.history, kHistoryBytes));
  HIPCHECK(hipMalloc(&R.history_count, sizeof(unsigned long long)));
  HIPCHECK(hipMemset(R.history_count, 0, sizeof(unsigned long long)));
  HIPCHECK(hipMalloc(&R.history_done, sizeof(unsigned int)));
  HIPCHECK(hipMemset(R.history_done, 0, sizeof(unsigned int)));
  HIPCHECK(hipStreamSynchronize(R.stream));
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal));
  R.capturing = true;
  return 0;
}

extern "C" int op2h_graph_end(void) {
  Runtime &R = rt();
  if (!R.capturing) return -1;
  op2h_flush_deferred();  // the last pending launch belongs to the graph
  size_t words = 0;
  HistSources src;
  std::memset(&src, 0, sizeof(src));
  if ((int)R.history_sources.size() > kHistSources) {
    printf("op2hip: more than %d gbl results in one captured graph\n",
           kHistSources);
    exit(-1);
  }
  for (auto &s : R.history_sources) {
    src.p[src.n] = (const unsigned long long *)s.first;
    src.words[src.n] = (int)((s.second + 7) / 8);
    src.off[src.n] = (int)words;
    src.kind[src.n] = 0;
    auto pi = R.partial_of_slot.find(s.first);
    if (pi != R.partial_of_slot.end()) {
      src.p[src.n] = (const unsigned long long *)pi->second.vals;
      src.locs[src.n] = pi->second.locs;
      src.kind[src.n] = pi->second.kind;
      src.nparts[src.n] = pi->second.nparts;
      src.red[src.nred++] = src.n;
    }
    words += src.words[src.n];
    src.n++;
  }
  R.history_words = (int)words;
  R.history_capacity =
      words ? std::max(1, (int)(kHistoryBytes / (words * sizeof(double)))) : 0;
  if (words)
    history_append<<<std::max(1, src.nred), op2h_red::kRedThreads, 0, R.stream>>>(
        src, R.history, R.history_count, R.history_done, R.history_capacity, (int)words);
  hipGraph_t g;
  HIPCHECK(hipStreamEndCapture(R.stream, &g));
  R.capturing = false;
  R.graph = g;
  HIPCHECK(hipGraphInstantiate(&R.graph_exec, g, nullptr, nullptr, 0));
  return 0;
}

extern "C" int op2h_graph_launch(int count) {
  Runtime &R = rt();
  if (!R.graph_exec || R.capturing) return -1;
  for (int i = 0; i < count; ++i) HIPCHECK(hipGraphLaunch(R.graph_exec, R.stream));
  return 0;
}

extern "C" int op2h_graph_sync(void) {
  Runtime &R = rt();
  if (R.capturing) return -1;
  HIPCHECK(hipStreamSynchronize(R.stream));
  // the latest launch's record holds every slot of the graph, in record order
  unsigned long long count = 0;
  if (R.history_count)
    HIPCHECK(hipMemcpy(&count, R.history_count, sizeof(count), hipMemcpyDeviceToHost));
  if (count == 0) return 0;  // graph never launched: host values stay as they are
  std::vector<double> rec(R.history_words);
  size_t pos = (size_t)((count - 1) % R.history_capacity) * R.history_words;
  HIPCHECK(hipMemcpy(rec.data(), R.history + pos, R.history_words * sizeof(double),
                     hipMemcpyDeviceToHost));
  size_t w = 0;
  for (auto &d : R.deferred) {
    std::memcpy(d.slot_h, &rec[w], d.bytes);
    w += (d.bytes + 7) / 8;
  }
  for (auto &d : R.deferred) {
    std::memcpy(d.host_ptr, d.snapshot.data(), d.snapshot.size());
    if (d.info_ptr)
      std::memcpy(d.info_ptr, d.info_snapshot.data(), d.info_snapshot.size());
    combine_gbl(d.acc, d.type, d.dim, d.host_ptr, d.slot_h, d.info_ptr,
                d.info_ptr ? d.slot_h + d.snapshot.size() : nullptr, d.info_size);
  }
  return 0;
}

extern "C" int op2h_graph_history(int max_records, double *out, int *nrecords) {
  Runtime &R = rt();
  if (nrecords) *nrecords = 0;
  if (!R.history || !R.history_words || R.capturing) return 0;
  HIPCHECK(hipStreamSynchronize(R.stream));
  unsigned long long count = 0;
  HIPCHECK(hipMemcpy(&count, R.history_count, sizeof(count), hipMemcpyDeviceToHost));
  unsigned long long avail = std::min<unsigned long long>(count, R.history_capacity);
  unsigned long long n = std::min<unsigned long long>(avail, (unsigned long long)std::max(max_records, 0));
  for (unsigned long long r = 0; r < n; ++r) {
    unsigned long long rec = count - n + r;
    size_t pos = (size_t)(rec % R.history_capacity) * R.history_words;
    HIPCHECK(hipMemcpy(out + r * R.history_words, R.history + pos,
                       R.history_words * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (nrecords) *nrecords = (int)n;
  return R.history_words;
}

extern "C" void op2h_graph_destroy(void) {
  Runtime &R = rt();
  if (!R.initialised) return;
  if (R.graph_exec || R.history) HIPCHECK(hipStreamSynchronize(R.stream));
  if (R.graph_exec) (void)hipGraphExecDestroy(R.graph_exec);
  if (R.graph) (void)hipGraphDestroy(R.graph);
  R.graph_exec = nullptr;
  R.graph = nullptr;
  for (auto &s : R.capture_slots) {
    (void)hipFree(s.first);
    (void)hipHostFree(s.second);
  }
  R.capture_slots.clear();
  R.capture_slot_next = 0;
  R.deferred.clear();
  R.history_sources.clear();
  R.partial_of_slot.clear();
  if (R.partial_pool) (void)hipFree(R.partial_pool);
  R.partial_pool = nullptr;
  R.partial_pool_bytes = R.partial_pool_next = 0;
  if (R.history) (void)hipFree(R.history);
  if (R.history_count) (void)hipFree(R.history_count);
  if (R.history_done) (void)hipFree(R.history_done);
  R.history_done = nullptr;
  R.history = nullptr;
  R.history_count = nullptr;
  R.history_words = 0;
  R.history_capacity = 0;
}
