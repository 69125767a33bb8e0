// airfoil_kernels.hip — the HIP-only target's output for the airfoil
// application (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:225-277): one
// op_par_loop_airfoil_<i>_<kernel> stub per loop, its gfx950 kernels and the
// application's device constants.
//
// User kernels (save_soln.h, adt_calc.h, res_calc.h, bres_calc.h, update.h) are
// device functions with the reference's expression trees; the whole library is
// compiled with -ffp-contract=off, so each multiply and add rounds as in the
// reference's host build.  Loops with indirect increments (res_calc,
// bres_calc) have four plans:
//   ATOMIC  one thread per edge, FP64 atomics on res;
//   COLOR   flat colouring, one launch per colour, plain read-modify-write;
//   STAGED  plan blocks staged in LDS, increments applied colour by colour
//           inside the block, blocks launched colour by colour;
//   GATHER  one thread per cell, the edges of the cell visited in edge order
//           and res accumulated in registers — the sequential executor's
//           summation order, so results match it bit for bit.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <tuple>
#include <type_traits>
#include <utility>

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstring>

#include <map>
#include <vector>

#include "op2_hip_airfoil.h"
#include "op2_hip_rt.h"
#include "op2h_plan_build.h"
#include "airfoil_device.h"
#include "op2h_reduce.h"

#define KCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

// The airfoil constants (airfoil.cpp:149-164), offered by op_decl_const_char.
static int airfoil_consts(int dim, char const *type, int size, const char *dat,
                          char const *name) {
  (void)type;
  size_t bytes = (size_t)dim * size;
  if (!strcmp(name, "gam")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(gam_d), dat, bytes));
  else if (!strcmp(name, "gm1")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(gm1_d), dat, bytes));
  else if (!strcmp(name, "cfl")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(cfl_d), dat, bytes));
  else if (!strcmp(name, "eps")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(eps_d), dat, bytes));
  else if (!strcmp(name, "mach")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(mach_d), dat, bytes));
  else if (!strcmp(name, "alpha")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(alpha_d), dat, bytes));
  else if (!strcmp(name, "qinf")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(qinf_d), dat, bytes));
  else return 0;
  // FP32 copies of the flux constants, rounded once on the host
  float f[4];
  for (int i = 0; i < dim && i < 4; ++i)
    f[i] = size == 8 ? (float)((const double *)dat)[i] : ((const float *)dat)[i];
  if (!strcmp(name, "gam")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(gam_f), f, sizeof(float)));
  else if (!strcmp(name, "gm1")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(gm1_f), f, sizeof(float)));
  else if (!strcmp(name, "cfl")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(cfl_f), f, sizeof(float)));
  else if (!strcmp(name, "eps")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(eps_f), f, sizeof(float)));
  else if (!strcmp(name, "qinf")) KCHECK(hipMemcpyToSymbol(HIP_SYMBOL(qinf_f), f, 4 * sizeof(float)));
  return 1;
}

// Arithmetic per element of each loop's kernel in multiply-add equivalents
// (a division or square root counted as ten), the cost model's measure of a
// loop's compute: save_soln copies; adt_calc has a division, five square
// roots and ~50 operations; res_calc two divisions and ~75 operations;
// bres_calc ~40; update a division and ~12.
static const struct AirfoilConstHook {
  AirfoilConstHook() {
    op2h_register_const_hook(airfoil_consts);
    op2h_set_loop_arith("save_soln", 4);
    op2h_set_loop_arith("adt_calc", 64);
    op2h_set_loop_arith("res_calc", 60);
    op2h_set_loop_arith("bres_calc", 40);
    op2h_set_loop_arith("update", 20);
  }
} airfoil_const_hook;


namespace {
using namespace op2_dev;

// Launch with the kernel's declared argument types; with timing events
// (op2h_kernel_timer_events) through hipExtLaunchKernel, which stamps `e0` at
// the kernel's start and `e1` at its end (either may be NULL).
template <class Tuple, std::size_t... I>
void arg_ptrs(Tuple &t, void **p, std::index_sequence<I...>) {
  ((p[I] = (void *)&std::get<I>(t)), ...);
}

template <typename... KArgs, typename... Args>
void launch_k(bool ext, void *e0, void *e1, void (*kernel)(KArgs...), dim3 grid,
              dim3 block, size_t lds, hipStream_t st, Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  std::tuple<KArgs...> t{static_cast<KArgs>(args)...};
  void *p[sizeof...(KArgs) > 0 ? sizeof...(KArgs) : 1];
  arg_ptrs(t, p, std::index_sequence_for<KArgs...>{});
  if (ext)
    KCHECK(hipExtLaunchKernel((const void *)kernel, grid, block, p, lds, st,
                              (hipEvent_t)e0, (hipEvent_t)e1, 0));
  else
    KCHECK(hipLaunchKernel((const void *)kernel, grid, block, p, lds, st));
}

// Timing for a loop launched with launch_k: kernel-attached events when
// timing is on, the event-pair timer (or the autotuner's) otherwise.
struct LoopTimer {
  op2h_loop_ctx *ctx;
  void *e0 = nullptr, *e1 = nullptr;
  bool ext = false;
  explicit LoopTimer(op2h_loop_ctx *c) : ctx(c) {
    ext = op2h_kernel_timer_events(c, &e0, &e1) != 0;
    if (!ext) op2h_kernel_timer_start(c);
  }
  void done() {
    if (!ext) op2h_kernel_timer_stop(ctx);
  }
};

// ---------------------------------------------------------------------------
// save_soln: qold = q over the whole dat, 16-byte vectors (direct copy)
// ---------------------------------------------------------------------------
__global__ void k_save_soln(const double2 *__restrict q, double2 *__restrict qold,
                            int n2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) qold[i] = q[i];
}

// ---------------------------------------------------------------------------
// adt_calc: one thread per cell, nodes gathered through pcell.  NT: the
// cell's own streams (its pcell row, q, adt) non-temporal, the gathered
// nodes (shared with the neighbouring cells) cached as usual.
// ---------------------------------------------------------------------------
template <bool SOA, bool NT>
__global__ void k_adt_calc(const double *__restrict x, const double *__restrict q,
                           double *__restrict adt, const int *__restrict pcell,
                           int n, int sx, int sq) {
  int c = xcd_block(kXcds) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int4 cn = NT ? ld_int4_nt(pcell, c) : reinterpret_cast<const int4 *>(pcell)[c];
  double x1[2], x2[2], x3[2], x4[2], qc[4], a;
  if (NT) ld4_nt<SOA>(q, c, sq, qc);
  else ld4<SOA>(q, c, sq, qc);
  ld2<SOA>(x, cn.x, sx, x1);
  ld2<SOA>(x, cn.y, sx, x2);
  ld2<SOA>(x, cn.z, sx, x3);
  ld2<SOA>(x, cn.w, sx, x4);
  op2_k::adt_calc(x1, x2, x3, x4, qc, &a);
  if (NT) __builtin_nontemporal_store(a, adt + c);
  else adt[c] = a;
}

// ---------------------------------------------------------------------------
// res_calc
// ---------------------------------------------------------------------------
template <bool SOA>
__device__ __forceinline__ void res_edge(const double *__restrict x,
                                         const double *__restrict q,
                                         const double *__restrict adt,
                                         int2 en, int2 ec, int sx, int sq,
                                         double *r1, double *r2) {
  double x1[2], x2[2], q1[4], q2[4];
  ld2<SOA>(x, en.x, sx, x1);
  ld2<SOA>(x, en.y, sx, x2);
  ld4<SOA>(q, ec.x, sq, q1);
  ld4<SOA>(q, ec.y, sq, q2);
  double a1 = adt[ec.x], a2 = adt[ec.y];
  for (int d = 0; d < 4; ++d) {
    r1[d] = 0.0;
    r2[d] = 0.0;
  }
  op2_k::res_calc(x1, x2, q1, q2, &a1, &a2, r1, r2);
}

template <bool SOA>
__global__ void k_res_atomic(const double *__restrict x, const double *__restrict q,
                             const double *__restrict adt, double *res,
                             const int2 *__restrict pedge,
                             const int2 *__restrict pecell, int n, int sx, int sq) {
  int e = xcd_block(kXcds) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int2 en = pedge[e], ec = pecell[e];
  double r1[4], r2[4];
  res_edge<SOA>(x, q, adt, en, ec, sx, sq, r1, r2);
  atomic_add4<SOA>(res, ec.x, sq, r1);
  atomic_add4<SOA>(res, ec.y, sq, r2);
}

template <bool SOA>
__global__ void k_res_color(const double *__restrict x, const double *__restrict q,
                            const double *__restrict adt, double *res,
                            const int2 *__restrict pedge,
                            const int2 *__restrict pecell,
                            const int *__restrict col_reord, int start, int end,
                            int sx, int sq) {
  int i = start + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= end) return;
  int e = col_reord[i];
  int2 en = pedge[e], ec = pecell[e];
  double r1[4], r2[4];
  res_edge<SOA>(x, q, adt, en, ec, sx, sq, r1, r2);
  add4<SOA>(res, ec.x, sq, r1);
  add4<SOA>(res, ec.y, sq, r2);
}

// Owner-compute: one thread per cell; entries are edge*2+slot in edge order.
template <bool SOA>
__global__ void k_res_gather(const double *__restrict x, const double *__restrict q,
                             const double *__restrict adt, double *res,
                             const int2 *__restrict pedge,
                             const int2 *__restrict pecell,
                             const int *__restrict offsets,
                             const int *__restrict entries,
                             const int *__restrict targets, int ntargets, int sx,
                             int sq) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntargets) return;
  int cell = targets[t];
  double acc[4];
  ld4<SOA>(res, cell, sq, acc);
  int k0 = offsets[t], k1 = offsets[t + 1];
  for (int k = k0; k < k1; ++k) {
    int ent = entries[k];
    int e = ent >> 1;
    int2 en = pedge[e], ec = pecell[e];
    double r1[4], r2[4];
    res_edge<SOA>(x, q, adt, en, ec, sx, sq, r1, r2);
    const double *r = (ent & 1) ? r2 : r1;
    acc[0] += r[0];
    acc[1] += r[1];
    acc[2] += r[2];
    acc[3] += r[3];
  }
  st4<SOA>(res, cell, sq, acc);
}

// Owner-compute with K lanes per cell: lane k of a cell's group computes the
// flux of the cell's k-th edge (in edge order); the group's first lane adds
// the K results to res in edge order, so the sum is the sequential
// executor's.  Every lane of a wave reaches every shuffle.
template <bool SOA, int K>
__global__ void __launch_bounds__(1024)
    k_res_gather_lanes(const double *__restrict x, const double *__restrict q,
                       const double *__restrict adt, double *res,
                       const int2 *__restrict pedge, const int2 *__restrict pecell,
                       const int *__restrict offsets, const int *__restrict entries,
                       const int *__restrict targets, int ntargets, int sx, int sq) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = g / K, k = g & (K - 1);
  const bool valid = t < ntargets;
  int k0 = 0, cnt = 0;
  if (valid) {
    k0 = offsets[t];
    cnt = offsets[t + 1] - k0;
  }
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (k < cnt) {
    int ent = entries[k0 + k];
    int e = ent >> 1;
    double r1[4], r2[4];
    res_edge<SOA>(x, q, adt, pedge[e], pecell[e], sx, sq, r1, r2);
    for (int d = 0; d < 4; ++d) r[d] = (ent & 1) ? r2[d] : r1[d];
  }
  const int lead = (threadIdx.x & 63) & ~(K - 1);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int cell = 0;
  if (k == 0 && valid) {
    cell = targets[t];
    ld4<SOA>(res, cell, sq, acc);
  }
  for (int j = 0; j < K; ++j) {
    double v[4];
    for (int d = 0; d < 4; ++d) v[d] = __shfl(r[d], lead + j, 64);
    if (j < cnt)
      for (int d = 0; d < 4; ++d) acc[d] += v[d];
  }
  if (k == 0 && valid) st4<SOA>(res, cell, sq, acc);
}

// Gather with a packed table: rec[t*K + k] = {n0, n1, other cell, slot} of
// the k-th edge of target t in edge order, {-1, -1, -1, 0} past its last edge.
template <int K>
__global__ void k_pack_res_gather(const int2 *__restrict pedge,
                                  const int2 *__restrict pecell,
                                  const int *__restrict offsets,
                                  const int *__restrict entries, int ntargets,
                                  int4 *__restrict rec) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ntargets * K) return;
  int t = (int)(i / K), k = (int)(i % K);
  int k0 = offsets[t], cnt = offsets[t + 1] - k0;
  int4 r = make_int4(-1, -1, -1, 0);
  if (k < cnt) {
    int ent = entries[k0 + k];
    int e = ent >> 1, slot = ent & 1;
    int2 en = pedge[e], ec = pecell[e];
    r = make_int4(en.x, en.y, slot ? ec.x : ec.y, slot);
  }
  rec[i] = r;
}

template <bool SOA, int K>
__global__ void __launch_bounds__(1024)
    k_res_gather_packed(const double *__restrict x, const double *__restrict q,
                        const double *__restrict adt, double *res,
                        const int4 *__restrict rec, const int *__restrict targets,
                        int ntargets, int sx, int sq) {
  const int g = xcd_block(kXcds) * blockDim.x + threadIdx.x;
  const int t = g / K, k = g & (K - 1);
  const bool valid = t < ntargets;
  const int cell = valid ? targets[t] : 0;
  const int4 rr = valid ? rec[g] : make_int4(-1, -1, -1, 0);
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  const int has = rr.x >= 0;
  if (has) {
    double x1[2], x2[2], qa[4], qb[4];
    ld2<SOA>(x, rr.x, sx, x1);
    ld2<SOA>(x, rr.y, sx, x2);
    ld4<SOA>(q, cell, sq, qa);
    ld4<SOA>(q, rr.z, sq, qb);
    // the edge's ecell[0] side is q1 whichever side this lane's cell is:
    // operands are selected, not branched on, so the wave evaluates one flux
    const bool second = rr.w != 0;
    double q1[4], q2[4];
    for (int d = 0; d < 4; ++d) {
      q1[d] = second ? qb[d] : qa[d];
      q2[d] = second ? qa[d] : qb[d];
    }
    const double aa = adt[cell], ab = adt[rr.z];
    double a1 = second ? ab : aa, a2 = second ? aa : ab;
    double r1[4] = {0.0, 0.0, 0.0, 0.0}, r2[4] = {0.0, 0.0, 0.0, 0.0};
    op2_k::res_calc(x1, x2, q1, q2, &a1, &a2, r1, r2);
    for (int d = 0; d < 4; ++d) r[d] = second ? r2[d] : r1[d];
  }
  static_assert(K == 4, "quad broadcasts assume four lanes per cell");
  // the quad's increments in edge order (lane 0 first), added to res
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (k == 0 && valid) ld4<SOA>(res, cell, sq, acc);
  double v[4][4];
  int hv[4];
  for (int d = 0; d < 4; ++d) {
    v[0][d] = quad_bcast<0>(r[d]);
    v[1][d] = quad_bcast<1>(r[d]);
    v[2][d] = quad_bcast<2>(r[d]);
    v[3][d] = quad_bcast<3>(r[d]);
  }
  hv[0] = quad_bcast<0>(has);
  hv[1] = quad_bcast<1>(has);
  hv[2] = quad_bcast<2>(has);
  hv[3] = quad_bcast<3>(has);
  for (int j = 0; j < 4; ++j)
    if (hv[j])
      for (int d = 0; d < 4; ++d) acc[d] += v[j][d];
  if (k == 0 && valid) st4<SOA>(res, cell, sq, acc);
}

struct StagedArgs {
  const int *blkmap, *offset, *nelems, *nthrcol;
  const unsigned char *thrcol;
  const int *ind_offs_n, *ind_sizes_n, *ind_map_n;  // node map slot
  const int *ind_offs_c, *ind_sizes_c, *ind_map_c;  // cell map slot
  const unsigned short *loc_n0, *loc_n1, *loc_c0, *loc_c1;
  int max_n, max_c;
};

// Two-level coloured, LDS-staged res_calc: one workgroup per plan block.
// LDS: x of the block's nodes [max_n][2], q [max_c][4], adt [max_c],
// res increments [max_c][4].
template <bool SOA>
__global__ void __launch_bounds__(1024)
    k_res_staged(const double *__restrict x, const double *__restrict q,
                 const double *__restrict adt, double *__restrict res,
                 StagedArgs P, int block_offset, int sx, int sq) {
  extern __shared__ double lds[];
  const int b = P.blkmap[blockIdx.x + block_offset];
  const int off = P.offset[b], nel = P.nelems[b];
  const int nn = P.ind_sizes_n[b], nbase = P.ind_offs_n[b];
  const int nc = P.ind_sizes_c[b], cbase = P.ind_offs_c[b];
  double *lx = lds;
  double *lq = lx + 2 * P.max_n;
  double *la = lq + 4 * P.max_c;
  double *lr = la + P.max_c;
  for (int k = threadIdx.x; k < nn; k += blockDim.x) {
    double v[2];
    ld2<SOA>(x, P.ind_map_n[nbase + k], sx, v);
    lx[2 * k] = v[0];
    lx[2 * k + 1] = v[1];
  }
  for (int k = threadIdx.x; k < nc; k += blockDim.x) {
    int cell = P.ind_map_c[cbase + k];
    double v[4];
    ld4<SOA>(q, cell, sq, v);
    lq[4 * k] = v[0];
    lq[4 * k + 1] = v[1];
    lq[4 * k + 2] = v[2];
    lq[4 * k + 3] = v[3];
    la[k] = adt[cell];
    lr[4 * k] = 0.0;
    lr[4 * k + 1] = 0.0;
    lr[4 * k + 2] = 0.0;
    lr[4 * k + 3] = 0.0;
  }
  __syncthreads();
  double r1[4] = {0.0, 0.0, 0.0, 0.0}, r2[4] = {0.0, 0.0, 0.0, 0.0};
  int col = -1, l0 = 0, l1 = 0;
  if ((int)threadIdx.x < nel) {
    const int e = off + threadIdx.x;
    const int a0 = P.loc_n0[e], a1 = P.loc_n1[e];
    l0 = P.loc_c0[e];
    l1 = P.loc_c1[e];
    op2_k::res_calc(lx + 2 * a0, lx + 2 * a1, lq + 4 * l0, lq + 4 * l1, la + l0,
                    la + l1, r1, r2);
    col = P.thrcol[e];
  }
  const int ncol = P.nthrcol[b];
  for (int c = 0; c < ncol; ++c) {
    if (col == c) {
      for (int d = 0; d < 4; ++d) {
        lr[4 * l0 + d] += r1[d];
        lr[4 * l1 + d] += r2[d];
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < nc; k += blockDim.x)
    add4<SOA>(res, P.ind_map_c[cbase + k], sq, lr + 4 * k);
}


// Per-instance records of res_calc's owner plan: {pedge[e], pecell[e]}.
__global__ void k_pack_owner_res(const int2 *__restrict pedge,
                                 const int2 *__restrict pecell,
                                 const int *__restrict elems, int total,
                                 int4 *__restrict inst) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  int e = elems[j];
  int2 n = pedge[e], c = pecell[e];
  inst[j] = make_int4(n.x, n.y, c.x, c.y);
}

// Compact records, one workgroup per block: the lowest cell and node its
// instances touch, and when every cell and node lies within 65535 of those,
// 8-byte records of 16-bit offsets (else the block keeps the 16-byte ones).
__global__ void __launch_bounds__(256)
    k_pack_owner_res8(const int4 *__restrict inst, const int *__restrict elem_offs,
                      int2 *__restrict base, uint2 *__restrict inst8) {
  __shared__ int s_min[2][4], s_max[2][4];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int e0 = elem_offs[b], e1 = elem_offs[b + 1];
  int cmin = INT_MAX, nmin = INT_MAX, cmax = -1, nmax = -1;
  for (int j = e0 + t; j < e1; j += blockDim.x) {
    const int4 r = inst[j];
    nmin = min(nmin, min(r.x, r.y));
    nmax = max(nmax, max(r.x, r.y));
    cmin = min(cmin, min(r.z, r.w));
    cmax = max(cmax, max(r.z, r.w));
  }
  for (int o = 32; o > 0; o >>= 1) {
    cmin = min(cmin, __shfl_xor(cmin, o));
    nmin = min(nmin, __shfl_xor(nmin, o));
    cmax = max(cmax, __shfl_xor(cmax, o));
    nmax = max(nmax, __shfl_xor(nmax, o));
  }
  if (lane == 0) {
    s_min[0][w] = cmin;
    s_min[1][w] = nmin;
    s_max[0][w] = cmax;
    s_max[1][w] = nmax;
  }
  __syncthreads();
  for (int k = 0; k < 4; ++k) {
    cmin = min(cmin, s_min[0][k]);
    nmin = min(nmin, s_min[1][k]);
    cmax = max(cmax, s_max[0][k]);
    nmax = max(nmax, s_max[1][k]);
  }
  const bool fits = e1 > e0 && cmax - cmin <= 0xFFFF && nmax - nmin <= 0xFFFF;
  if (t == 0) base[b] = fits ? make_int2(cmin, nmin) : make_int2(0, -1);
  if (!fits) return;
  for (int j = e0 + t; j < e1; j += blockDim.x) {
    const int4 r = inst[j];
    inst8[j] = make_uint2((unsigned)(r.x - nmin) | ((unsigned)(r.y - nmin) << 16),
                          (unsigned)(r.z - cmin) | ((unsigned)(r.w - cmin) << 16));
  }
}

// Owner plan for res_calc's args (res incremented via pecell 0 (+f) and 1
// (-f)), with its instance records packed once per plan; a NULL plan or an
// unexpected shape is fatal, as the reference's plan errors are.
static OwnerArgs owner_args_res(const op2h_owner_plan *p, const op_arg *args,
                                hipStream_t st) {
  if (!p) exit(-1);
  if (p->inc_of_arg[6] != 0 || p->inc_of_arg[7] != 1 || p->ninc != 2) {
    printf("op2hip: res_calc owner plan expects res via ecell[0], ecell[1]\n");
    exit(-1);
  }
  // packed once per plan (the plan pins the set, the maps and the block size)
  char key[96];
  snprintf(key, sizeof key, "res_calc.owner.%p", (const void *)p);
  int fresh = 0;
  int4 *inst = (int4 *)op2h_plan_user(key, sizeof(int4) * (size_t)p->total, &fresh);
  if (fresh && p->total)
    k_pack_owner_res<<<grid_for(p->total, 256), 256, 0, st>>>(
        (const int2 *)args[0].map->map_d, (const int2 *)args[2].map->map_d, p->elems,
        p->total, inst);
  // compact records: blocks per launch are at most a few hundred thousand, so
  // one workgroup per block
  snprintf(key, sizeof key, "res_calc.owner8.%p", (const void *)p);
  const size_t base_bytes = (sizeof(int2) * (size_t)p->nblocks + 255) / 256 * 256;
  char *c8 = (char *)op2h_plan_user(key, base_bytes + sizeof(uint2) * (size_t)p->total, &fresh);
  if (fresh && p->nblocks)
    k_pack_owner_res8<<<p->nblocks, 256, 0, st>>>(inst, p->elem_offs, (int2 *)c8,
                                                  (uint2 *)(c8 + base_bytes));
  OwnerArgs A;
  A.inst = inst;
  A.base = (const int2 *)c8;
  A.inst8 = (const uint2 *)(c8 + base_bytes);
  A.elem_offs = p->elem_offs;
  A.owned_offs = p->owned_offs;
  A.owned = p->owned;
  A.csr_offs = p->csr_offs;
  A.csr = p->csr;
  // the plan's fixed-width increment lists (quadrilateral cells: at most
  // four edges each): one 8-byte load per owned cell
  A.csr4 = (const ushort4 *)p->csr4;
  A.max_e = p->max_elems;
  return A;
}

// Owner-compute res_calc (op2_dev::owner_res_block), one workgroup per block
// of owned cells, blocks in XCD order.  LDS: fluxes [max_e][4].
template <bool SOA, bool C4>
__global__ void __launch_bounds__(1024)
    k_res_owner(const double *__restrict x, const double *__restrict q,
                const double *__restrict adt, double *__restrict res,
                OwnerArgs P, int sx, int sq, int zero_acc) {
  extern __shared__ double lf[];
  owner_res_block<SOA, C4>(xcd_block(kXcds), x, q, adt, res, P, sx, sq, lf, zero_acc != 0);
}

// ---------------------------------------------------------------------------
// bres_calc
// ---------------------------------------------------------------------------
template <bool SOA>
__device__ __forceinline__ void bres_edge(const double *__restrict x,
                                          const double *__restrict q,
                                          const double *__restrict adt,
                                          int2 bn, int cell, int bound, int sx,
                                          int sq, double *r) {
  double x1[2], x2[2], q1[4];
  ld2<SOA>(x, bn.x, sx, x1);
  ld2<SOA>(x, bn.y, sx, x2);
  ld4<SOA>(q, cell, sq, q1);
  double a1 = adt[cell];
  for (int d = 0; d < 4; ++d) r[d] = 0.0;
  op2_k::bres_calc(x1, x2, q1, &a1, r, &bound);
}

template <bool SOA>
__global__ void k_bres_atomic(const double *__restrict x, const double *__restrict q,
                              const double *__restrict adt, double *res,
                              const int2 *__restrict pbedge,
                              const int *__restrict pbecell,
                              const int *__restrict bound, int n, int sx, int sq) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int cell = pbecell[e];
  double r[4];
  bres_edge<SOA>(x, q, adt, pbedge[e], cell, bound[e], sx, sq, r);
  atomic_add4<SOA>(res, cell, sq, r);
}

template <bool SOA>
__global__ void k_bres_color(const double *__restrict x, const double *__restrict q,
                             const double *__restrict adt, double *res,
                             const int2 *__restrict pbedge,
                             const int *__restrict pbecell,
                             const int *__restrict bound,
                             const int *__restrict col_reord, int start, int end,
                             int sx, int sq) {
  int i = start + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= end) return;
  int e = col_reord[i];
  int cell = pbecell[e];
  double r[4];
  bres_edge<SOA>(x, q, adt, pbedge[e], cell, bound[e], sx, sq, r);
  add4<SOA>(res, cell, sq, r);
}

template <bool SOA>
__global__ void k_bres_gather(const double *__restrict x, const double *__restrict q,
                              const double *__restrict adt, double *res,
                              const int2 *__restrict pbedge,
                              const int *__restrict bound,
                              const int *__restrict offsets,
                              const int *__restrict entries,
                              const int *__restrict targets, int ntargets, int sx,
                              int sq) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntargets) return;
  int cell = targets[t];
  double acc[4];
  ld4<SOA>(res, cell, sq, acc);
  for (int k = offsets[t]; k < offsets[t + 1]; ++k) {
    int e = entries[k];  // map dim 1: entry = bedge
    double r[4];
    bres_edge<SOA>(x, q, adt, pbedge[e], cell, bound[e], sx, sq, r);
    acc[0] += r[0];
    acc[1] += r[1];
    acc[2] += r[2];
    acc[3] += r[3];
  }
  st4<SOA>(res, cell, sq, acc);
}

struct BStagedArgs {
  const int *blkmap, *offset, *nelems, *nthrcol;
  const unsigned char *thrcol;
  const int *ind_offs_n, *ind_sizes_n, *ind_map_n;
  const int *ind_offs_c, *ind_sizes_c, *ind_map_c;
  const unsigned short *loc_n0, *loc_n1, *loc_c0;
  int max_n, max_c;
};

template <bool SOA>
__global__ void __launch_bounds__(1024)
    k_bres_staged(const double *__restrict x, const double *__restrict q,
                  const double *__restrict adt, double *__restrict res,
                  const int *__restrict bound, BStagedArgs P, int block_offset,
                  int sx, int sq) {
  extern __shared__ double lds[];
  const int b = P.blkmap[blockIdx.x + block_offset];
  const int off = P.offset[b], nel = P.nelems[b];
  const int nn = P.ind_sizes_n[b], nbase = P.ind_offs_n[b];
  const int nc = P.ind_sizes_c[b], cbase = P.ind_offs_c[b];
  double *lx = lds;
  double *lq = lx + 2 * P.max_n;
  double *la = lq + 4 * P.max_c;
  double *lr = la + P.max_c;
  for (int k = threadIdx.x; k < nn; k += blockDim.x) {
    double v[2];
    ld2<SOA>(x, P.ind_map_n[nbase + k], sx, v);
    lx[2 * k] = v[0];
    lx[2 * k + 1] = v[1];
  }
  for (int k = threadIdx.x; k < nc; k += blockDim.x) {
    int cell = P.ind_map_c[cbase + k];
    double v[4];
    ld4<SOA>(q, cell, sq, v);
    for (int d = 0; d < 4; ++d) {
      lq[4 * k + d] = v[d];
      lr[4 * k + d] = 0.0;
    }
    la[k] = adt[cell];
  }
  __syncthreads();
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  int col = -1, l0 = 0;
  if ((int)threadIdx.x < nel) {
    const int e = off + threadIdx.x;
    const int a0 = P.loc_n0[e], a1 = P.loc_n1[e];
    l0 = P.loc_c0[e];
    int bd = bound[e];
    op2_k::bres_calc(lx + 2 * a0, lx + 2 * a1, lq + 4 * l0, la + l0, r, &bd);
    col = P.thrcol[e];
  }
  const int ncol = P.nthrcol[b];
  for (int c = 0; c < ncol; ++c) {
    if (col == c)
      for (int d = 0; d < 4; ++d) lr[4 * l0 + d] += r[d];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < nc; k += blockDim.x)
    add4<SOA>(res, P.ind_map_c[cbase + k], sq, lr + 4 * k);
}

// ---------------------------------------------------------------------------
// update, with the rms (INC) and maxerr (MAX, errloc argmax) reductions.
// Each workgroup reduces its elements (waves by shuffles, then LDS) and writes
// one partial; a one-workgroup kernel then combines the partials in
// workgroup order and writes the result slots.  The kernel boundary makes the
// partials visible across XCDs, so no device-scope fence is needed inside the
// streaming kernel (on a multi-XCD gfx950 such a fence writes back L2).
// ---------------------------------------------------------------------------
struct MaxLoc {
  double v;
  int loc;
};

__device__ __forceinline__ MaxLoc maxloc_combine(MaxLoc a, MaxLoc b) {
  // larger value wins; equal values keep the lower element index, so the
  // result is the first element in set order reaching the maximum
  if (b.v > a.v || (b.v == a.v && b.loc < a.loc)) return b;
  return a;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ MaxLoc wave_maxloc(MaxLoc m) {
  for (int o = 32; o > 0; o >>= 1) {
    MaxLoc t;
    t.v = __shfl_xor(m.v, o, 64);
    t.loc = __shfl_xor(m.loc, o, 64);
    m = maxloc_combine(m, t);
  }
  return m;
}

// The cells a workgroup of the update (or of a direct chain fused with it)
// takes, and the partial it writes: with chunk > 0 the contiguous cells
// [L chunk, (L+1) chunk) of the workgroup the XCD-ordered id L names, its
// threads striding by the workgroup size; with chunk 0 a grid-stride sweep
// from the workgroup's own id.  The same cells in the same order per
// partial, so every kernel launched with one shape reduces the same bits.
struct CellRange {
  int L, c, end, step;
};
__device__ __forceinline__ CellRange cell_range(int n, int chunk) {
  if (chunk > 0) {
    const int L = xcd_block(kXcds);
    return {L, L * chunk + (int)threadIdx.x, min(n, (L + 1) * chunk), (int)blockDim.x};
  }
  return {(int)blockIdx.x, (int)(blockIdx.x * blockDim.x + threadIdx.x), n,
          (int)(gridDim.x * blockDim.x)};
}

template <bool SOA>
__global__ void __launch_bounds__(1024)
    k_update(const double *__restrict qold, double *__restrict q,
             double *__restrict res, const double *__restrict adt, int n, int s,
             double *__restrict part_sum, double *__restrict part_max,
             int *__restrict part_loc, const int *__restrict oldidx, int chunk) {
  __shared__ double s_sum[16];
  __shared__ double s_max[16];
  __shared__ int s_loc[16];
  double rms = 0.0;
  MaxLoc ml{-INFINITY, 0x7fffffff};
  const CellRange cr = cell_range(n, chunk);
  for (int c = cr.c; c < cr.end; c += cr.step) {
    double qo[4], qn[4], r[4], a = adt[c], mx = ml.v;
    int loc = ml.loc;
    ld4<SOA>(qold, c, s, qo);
    ld4<SOA>(res, c, s, r);
    // op_arg_idx: the application's element index (transparent renumbering)
    int ci = oldidx ? oldidx[c] : c;
    op2_k::update(qo, qn, r, &a, &rms, &mx, &ci, &loc);
    ml.v = mx;
    ml.loc = loc;
    st4<SOA>(q, c, s, qn);
    st4<SOA>(res, c, s, r);
  }
  rms = wave_sum(rms);
  ml = wave_maxloc(ml);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_sum[wave] = rms;
    s_max[wave] = ml.v;
    s_loc[wave] = ml.loc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bs = 0.0;
    MaxLoc bm{-INFINITY, 0x7fffffff};
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) {
      bs += s_sum[w];
      bm = maxloc_combine(bm, MaxLoc{s_max[w], s_loc[w]});
    }
    part_sum[cr.L] = bs;
    part_max[cr.L] = bm.v;
    part_loc[cr.L] = bm.loc;
  }
}

// Two groups of 256 threads: group 0 finishes rms (sum), group 1 maxerr and
// errloc (max, lowest element on ties), both with op2h_red::group_reduce, the
// reduction the graph's history kernel applies to deferred partials.
constexpr int kFinishThreads = 2 * op2h_red::kRedThreads;

__global__ void __launch_bounds__(kFinishThreads)
    k_update_finish(const double *__restrict part_sum,
                    const double *__restrict part_max,
                    const int *__restrict part_loc, int nparts,
                    double *__restrict slot_rms, double *__restrict slot_max) {
  __shared__ double s_v[2][op2h_red::kRedThreads / 64];
  __shared__ int s_l[2][op2h_red::kRedThreads / 64];
  const int grp = threadIdx.x / op2h_red::kRedThreads;
  const int t = threadIdx.x % op2h_red::kRedThreads;
  // one call site for both groups (the reduction holds a workgroup barrier)
  const op2h_red::ValLoc r =
      op2h_red::group_reduce(grp == 0 ? part_sum : part_max, grp == 0 ? nullptr : part_loc,
                             nparts, grp == 0 ? 1 : 2, t, s_v[grp], s_l[grp]);
  if (t == 0) {
    if (grp == 0) {
      slot_rms[0] = r.v;
    } else {
      slot_max[0] = r.v;
      reinterpret_cast<int *>(slot_max + 1)[0] = r.loc;
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed precision: adt and res stored in FP32, fluxes and adt computed in FP32
// from FP32 copies of q and x; q, qold, the update and the rms/maxerr
// reductions stay FP64.
// ---------------------------------------------------------------------------
template <bool SOA>
__device__ __forceinline__ void ld2f(const double *__restrict p, int n, int s, float *v) {
  double t[2];
  ld2<SOA>(p, n, s, t);
  v[0] = (float)t[0];
  v[1] = (float)t[1];
}

template <bool SOA>
__device__ __forceinline__ void ld4f(const double *__restrict p, int n, int s, float *v) {
  double t[4];
  ld4<SOA>(p, n, s, t);
  for (int d = 0; d < 4; ++d) v[d] = (float)t[d];
}

template <bool SOA>
__device__ __forceinline__ void ld4s(const float *__restrict p, int n, int s, float *v) {
  if (SOA) {
    for (int d = 0; d < 4; ++d) v[d] = p[(size_t)d * s + n];
  } else {
    float4 a = *reinterpret_cast<const float4 *>(p + 4 * (size_t)n);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
  }
}

template <bool SOA>
__device__ __forceinline__ void st4s(float *__restrict p, int n, int s, const float *v) {
  if (SOA) {
    for (int d = 0; d < 4; ++d) p[(size_t)d * s + n] = v[d];
  } else {
    *reinterpret_cast<float4 *>(p + 4 * (size_t)n) = make_float4(v[0], v[1], v[2], v[3]);
  }
}

template <bool SOA>
__device__ __forceinline__ void atomic_add4s(float *p, int n, int s, const float *v) {
  size_t st = SOA ? (size_t)s : 1, base = SOA ? (size_t)n : 4 * (size_t)n;
  for (int d = 0; d < 4; ++d) unsafeAtomicAdd(p + base + d * st, v[d]);
}

template <bool SOA>
__global__ void k_adt_calc_mixed(const double *__restrict x, const double *__restrict q,
                                 float *__restrict adt, const int *__restrict pcell,
                                 int n, int sx, int sq) {
  int c = xcd_block(kXcds) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int4 cn = reinterpret_cast<const int4 *>(pcell)[c];
  float x1[2], x2[2], x3[2], x4[2], qc[4], a;
  ld2f<SOA>(x, cn.x, sx, x1);
  ld2f<SOA>(x, cn.y, sx, x2);
  ld2f<SOA>(x, cn.z, sx, x3);
  ld2f<SOA>(x, cn.w, sx, x4);
  ld4f<SOA>(q, c, sq, qc);
  op2_kf::adt_calc(x1, x2, x3, x4, qc, &a);
  adt[c] = a;
}

template <bool SOA>
__global__ void k_res_atomic_mixed(const double *__restrict x, const double *__restrict q,
                                   const float *__restrict adt, float *res,
                                   const int2 *__restrict pedge,
                                   const int2 *__restrict pecell, int n, int sx, int sq) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int2 en = pedge[e], ec = pecell[e];
  float x1[2], x2[2], q1[4], q2[4];
  ld2f<SOA>(x, en.x, sx, x1);
  ld2f<SOA>(x, en.y, sx, x2);
  ld4f<SOA>(q, ec.x, sq, q1);
  ld4f<SOA>(q, ec.y, sq, q2);
  float a1 = adt[ec.x], a2 = adt[ec.y];
  float r1[4] = {0.f, 0.f, 0.f, 0.f}, r2[4] = {0.f, 0.f, 0.f, 0.f};
  op2_kf::res_calc(x1, x2, q1, q2, &a1, &a2, r1, r2);
  atomic_add4s<SOA>(res, ec.x, sq, r1);
  atomic_add4s<SOA>(res, ec.y, sq, r2);
}

template <bool SOA, int K>
__global__ void __launch_bounds__(1024)
    k_res_gather_packed_mixed(const double *__restrict x, const double *__restrict q,
                              const float *__restrict adt, float *res,
                              const int4 *__restrict rec, const int *__restrict targets,
                              int ntargets, int sx, int sq) {
  const int g = xcd_block(kXcds) * blockDim.x + threadIdx.x;
  const int t = g / K, k = g & (K - 1);
  const bool valid = t < ntargets;
  const int cell = valid ? targets[t] : 0;
  const int4 rr = valid ? rec[g] : make_int4(-1, -1, -1, 0);
  float r[4] = {0.f, 0.f, 0.f, 0.f};
  const int has = rr.x >= 0;
  if (has) {
    float x1[2], x2[2], qa[4], qb[4];
    ld2f<SOA>(x, rr.x, sx, x1);
    ld2f<SOA>(x, rr.y, sx, x2);
    ld4f<SOA>(q, cell, sq, qa);
    ld4f<SOA>(q, rr.z, sq, qb);
    const bool second = rr.w != 0;
    float q1[4], q2[4];
    for (int d = 0; d < 4; ++d) {
      q1[d] = second ? qb[d] : qa[d];
      q2[d] = second ? qa[d] : qb[d];
    }
    const float aa = adt[cell], ab = adt[rr.z];
    float a1 = second ? ab : aa, a2 = second ? aa : ab;
    float r1[4] = {0.f, 0.f, 0.f, 0.f}, r2[4] = {0.f, 0.f, 0.f, 0.f};
    op2_kf::res_calc(x1, x2, q1, q2, &a1, &a2, r1, r2);
    for (int d = 0; d < 4; ++d) r[d] = second ? r2[d] : r1[d];
  }
  static_assert(K == 4, "quad broadcasts assume four lanes per cell");
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  if (k == 0 && valid) ld4s<SOA>(res, cell, sq, acc);
  float v[4][4];
  int hv[4];
  for (int d = 0; d < 4; ++d) {
    v[0][d] = quad_bcast<0>(r[d]);
    v[1][d] = quad_bcast<1>(r[d]);
    v[2][d] = quad_bcast<2>(r[d]);
    v[3][d] = quad_bcast<3>(r[d]);
  }
  hv[0] = quad_bcast<0>(has);
  hv[1] = quad_bcast<1>(has);
  hv[2] = quad_bcast<2>(has);
  hv[3] = quad_bcast<3>(has);
  for (int j = 0; j < 4; ++j)
    if (hv[j])
      for (int d = 0; d < 4; ++d) acc[d] += v[j][d];
  if (k == 0 && valid) st4s<SOA>(res, cell, sq, acc);
}

// Mixed precision (SP fluxes, FP32 adt and res; x and q read as FP64 and
// rounded), same structure as k_res_owner.
template <bool SOA, bool C4>
__global__ void __launch_bounds__(1024)
    k_res_owner_mixed(const double *__restrict x, const double *__restrict q,
                      const float *__restrict adt, float *__restrict res,
                      OwnerArgs P, int sx, int sq, int zero_acc) {
  extern __shared__ float lff[];
  const int b = xcd_block(kXcds);
  const int e0 = P.elem_offs[b], ne = P.elem_offs[b + 1] - e0;
  const int2 bb = P.base[b];
  for (int j = threadIdx.x; j < ne; j += blockDim.x) {
    const int4 r = owner_record(P, bb, e0 + j);
    float x1[2], x2[2], q1[4], q2[4];
    ld2f<SOA>(x, r.x, sx, x1);
    ld2f<SOA>(x, r.y, sx, x2);
    ld4f<SOA>(q, r.z, sq, q1);
    ld4f<SOA>(q, r.w, sq, q2);
    op2_kf::res_flux(x1, x2, q1, q2, adt[r.z], adt[r.w], lff + 4 * j);
  }
  __syncthreads();
  const int o0 = P.owned_offs[b], no = P.owned_offs[b + 1] - o0;
  for (int i = threadIdx.x; i < no; i += blockDim.x) {
    const int cell = P.owned[o0 + i];
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    if (!zero_acc) ld4s<SOA>(res, cell, sq, acc);
    if (C4) {
      const ushort4 e = P.csr4[o0 + i];
      if (e.x != 0xFFFF) owner_apply(acc, lff, e.x);
      if (e.y != 0xFFFF) owner_apply(acc, lff, e.y);
      if (e.z != 0xFFFF) owner_apply(acc, lff, e.z);
      if (e.w != 0xFFFF) owner_apply(acc, lff, e.w);
    } else {
      for (int k = P.csr_offs[o0 + i]; k < P.csr_offs[o0 + i + 1]; ++k)
        owner_apply(acc, lff, (int)P.csr[k]);
    }
    st4s<SOA>(res, cell, sq, acc);
  }
}

template <bool SOA>
__device__ __forceinline__ void bres_edge_mixed(const double *__restrict x,
                                                const double *__restrict q,
                                                const float *__restrict adt, int2 bn,
                                                int cell, int bound, int sx, int sq,
                                                float *r) {
  float x1[2], x2[2], q1[4];
  ld2f<SOA>(x, bn.x, sx, x1);
  ld2f<SOA>(x, bn.y, sx, x2);
  ld4f<SOA>(q, cell, sq, q1);
  float a1 = adt[cell];
  for (int d = 0; d < 4; ++d) r[d] = 0.f;
  op2_kf::bres_calc(x1, x2, q1, &a1, r, &bound);
}

template <bool SOA>
__global__ void k_bres_atomic_mixed(const double *__restrict x, const double *__restrict q,
                                    const float *__restrict adt, float *res,
                                    const int2 *__restrict pbedge,
                                    const int *__restrict pbecell,
                                    const int *__restrict bound, int n, int sx, int sq) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int cell = pbecell[e];
  float r[4];
  bres_edge_mixed<SOA>(x, q, adt, pbedge[e], cell, bound[e], sx, sq, r);
  atomic_add4s<SOA>(res, cell, sq, r);
}

template <bool SOA>
__global__ void k_bres_gather_mixed(const double *__restrict x, const double *__restrict q,
                                    const float *__restrict adt, float *res,
                                    const int2 *__restrict pbedge,
                                    const int *__restrict bound,
                                    const int *__restrict offsets,
                                    const int *__restrict entries,
                                    const int *__restrict targets, int ntargets, int sx,
                                    int sq) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntargets) return;
  int cell = targets[t];
  float acc[4];
  ld4s<SOA>(res, cell, sq, acc);
  for (int k = offsets[t]; k < offsets[t + 1]; ++k) {
    int e = entries[k];
    float r[4];
    bres_edge_mixed<SOA>(x, q, adt, pbedge[e], cell, bound[e], sx, sq, r);
    for (int d = 0; d < 4; ++d) acc[d] += r[d];
  }
  st4s<SOA>(res, cell, sq, acc);
}

template <bool SOA>
__global__ void __launch_bounds__(1024)
    k_update_mixed(const double *__restrict qold, double *__restrict q,
                   float *__restrict res, const float *__restrict adt, int n, int s,
                   double *__restrict part_sum, double *__restrict part_max,
                   int *__restrict part_loc, const int *__restrict oldidx, int chunk) {
  __shared__ double s_sum[16];
  __shared__ double s_max[16];
  __shared__ int s_loc[16];
  double rms = 0.0;
  MaxLoc ml{-INFINITY, 0x7fffffff};
  const CellRange cr = cell_range(n, chunk);
  for (int c = cr.c; c < cr.end; c += cr.step) {
    double qo[4], qn[4], r[4], a = (double)adt[c], mx = ml.v;
    float rf[4];
    int loc = ml.loc;
    ld4<SOA>(qold, c, s, qo);
    ld4s<SOA>(res, c, s, rf);
    for (int d = 0; d < 4; ++d) r[d] = (double)rf[d];
    // op_arg_idx: the application's element index (transparent renumbering)
    int ci = oldidx ? oldidx[c] : c;
    op2_k::update(qo, qn, r, &a, &rms, &mx, &ci, &loc);
    ml.v = mx;
    ml.loc = loc;
    st4<SOA>(q, c, s, qn);
    const float z[4] = {0.f, 0.f, 0.f, 0.f};
    st4s<SOA>(res, c, s, z);
  }
  rms = wave_sum(rms);
  ml = wave_maxloc(ml);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_sum[wave] = rms;
    s_max[wave] = ml.v;
    s_loc[wave] = ml.loc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bs = 0.0;
    MaxLoc bm{-INFINITY, 0x7fffffff};
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) {
      bs += s_sum[w];
      bm = maxloc_combine(bm, MaxLoc{s_max[w], s_loc[w]});
    }
    part_sum[cr.L] = bs;
    part_max[cr.L] = bm.v;
    part_loc[cr.L] = bm.loc;
  }
}

}  // namespace

// ===========================================================================
// loop stubs
// ===========================================================================
namespace {

bool loop_soa(const op_arg *args, int nargs, int *stride) {
  int soa = -1;
  *stride = 0;
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || !a.opt || a.dat->dim <= 1) continue;
    int s = op2h_dat_soa_stride(a.dat);
    int this_soa = s > 0;
    if (soa >= 0 && this_soa != soa) {
      printf("op2hip: loop mixes AoS and SoA dats; declare all dats of a loop "
             "with the same layout\n");
      exit(-1);
    }
    soa = this_soa;
  }
  return soa == 1;
}

bool is_float(const op_arg &a) { return a.dat && !strncmp(a.dat->type, "float", 5); }

int set_stride(const op_arg &a) {
  int s = op2h_dat_soa_stride(a.dat);
  return s;
}

void account(int kernel_idx, op_set set, int nargs, const op_arg *args) {
  OP_kernels[kernel_idx].transfer += (float)op2h_algorithmic_bytes(set, nargs, args);
}

}  // namespace

// ---------------------------------------------------------------------------
// Loop-chain fusion inside a captured step.  The stubs leave launches pending
// (op2h_defer) and the next stub folds them into its own:
//   - the direct loops of one cell set run per cell in one kernel: an update
//     (and/or a save_soln) is held until the adt_calc that follows and then
//     launched as [update] [save_soln] adt_calc per cell (adt depends on its
//     own cell's q only);
//   - an owner-plan res_calc is held until the bres_calc that follows and
//     both run as one owner launch, each owned cell adding its boundary
//     fluxes after its interior ones.
// Every cell sees the operations of the unfused loops in the same order, and
// the update keeps its grid, block and partials, so every result is the
// unfused one bit for bit.  Anything else flushes a pending launch as it was.
// ---------------------------------------------------------------------------
namespace {

// R: the type of res and adt, double, or float in mixed mode (the update in
// FP64 from FP32 res and adt, adt_calc in FP32 from x and q rounded, as
// k_update_mixed and k_adt_calc_mixed).
template <class R, bool SOA, bool UPD, bool SAVE>
__global__ void __launch_bounds__(1024)
    k_direct_chain(const double *qold, double *__restrict q, R *__restrict res,
                   R *adt, double *qold_out, const double *__restrict x,
                   const int *__restrict pcell, const uint2 *__restrict prow,
                   const int *__restrict pbase, int n, int s, int sx,
                   double *__restrict part_sum, double *__restrict part_max,
                   int *__restrict part_loc, const int *__restrict oldidx, int chunk,
                   int res_dead) {
  // qold (read by update, written by save_soln) and adt (read by update,
  // written by adt_calc) are read before they are written, by the same
  // thread for the same cell: not restrict-qualified.  res_dead: the update's
  // zeros in res are not stored (the res_calc that follows in the same step
  // starts its sums from +0.0 instead of reading them)
  constexpr bool MIX = std::is_same<R, float>::value;
  __shared__ double s_sum[16];
  __shared__ double s_max[16];
  __shared__ int s_loc[16];
  double rms = 0.0;
  MaxLoc ml{-INFINITY, 0x7fffffff};
  // the update's cells per workgroup (its partials are the update's)
  const CellRange cr = cell_range(n, chunk);
  for (int c = cr.c; c < cr.end; c += cr.step) {
    double qn[4];
    if (UPD) {
      double qo[4], r[4], a = (double)adt[c], mx = ml.v;
      int loc = ml.loc;
      ld4<SOA>(qold, c, s, qo);
      if constexpr (MIX) {
        float rf[4];
        ld4s<SOA>(res, c, s, rf);
        for (int d = 0; d < 4; ++d) r[d] = (double)rf[d];
      } else {
        ld4<SOA>(res, c, s, r);
      }
      int ci = oldidx ? oldidx[c] : c;
      op2_k::update(qo, qn, r, &a, &rms, &mx, &ci, &loc);
      ml.v = mx;
      ml.loc = loc;
      st4<SOA>(q, c, s, qn);
      if (!res_dead) {
        if constexpr (MIX) {
          const float z[4] = {0.f, 0.f, 0.f, 0.f};
          st4s<SOA>(res, c, s, z);
        } else {
          st4<SOA>(res, c, s, r);
        }
      }
    } else {
      ld4<SOA>(q, c, s, qn);
    }
    if (SAVE) st4<SOA>(qold_out, c, s, qn);
    int4 cn;
    if (prow) {  // compact rows: four 16-bit offsets from the cell group's base
      const uint2 u = prow[c];
      const int b = pbase[c >> 6];
      cn = make_int4(b + (int)(u.x & 0xFFFFu), b + (int)(u.x >> 16), b + (int)(u.y & 0xFFFFu),
                     b + (int)(u.y >> 16));
    } else {
      cn = reinterpret_cast<const int4 *>(pcell)[c];
    }
    if constexpr (MIX) {
      float x1[2], x2[2], x3[2], x4[2], qf[4], a2;
      ld2f<SOA>(x, cn.x, sx, x1);
      ld2f<SOA>(x, cn.y, sx, x2);
      ld2f<SOA>(x, cn.z, sx, x3);
      ld2f<SOA>(x, cn.w, sx, x4);
      for (int d = 0; d < 4; ++d) qf[d] = (float)qn[d];
      op2_kf::adt_calc(x1, x2, x3, x4, qf, &a2);
      adt[c] = a2;
    } else {
      double x1[2], x2[2], x3[2], x4[2], a2;
      ld2<SOA>(x, cn.x, sx, x1);
      ld2<SOA>(x, cn.y, sx, x2);
      ld2<SOA>(x, cn.z, sx, x3);
      ld2<SOA>(x, cn.w, sx, x4);
      op2_k::adt_calc(x1, x2, x3, x4, qn, &a2);
      adt[c] = a2;
    }
  }
  if (!UPD) return;
  // the partials of k_update, by the same tree
  rms = wave_sum(rms);
  ml = wave_maxloc(ml);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_sum[wave] = rms;
    s_max[wave] = ml.v;
    s_loc[wave] = ml.loc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bsum = 0.0;
    MaxLoc bm{-INFINITY, 0x7fffffff};
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) {
      bsum += s_sum[w];
      bm = maxloc_combine(bm, MaxLoc{s_max[w], s_loc[w]});
    }
    part_sum[cr.L] = bsum;
    part_max[cr.L] = bm.v;
    part_loc[cr.L] = bm.loc;
  }
}

template <bool SOA>
__global__ void __launch_bounds__(1024)
    k_res_bres_owner(const double *__restrict x, const double *__restrict q,
                     const double *__restrict adt, double *__restrict res, OwnerArgs P,
                     BoundaryArgs B, int sx, int sq, int zero_acc) {
  extern __shared__ double lf[];
  owner_res_bres_block<SOA>(xcd_block(kXcds), x, q, adt, res, P, B, sx, sq, lf, zero_acc != 0);
}

// A pending run of direct loops over one cell set: [update] [save_soln].
// The update's launch shape: g workgroups, one reduction partial each
// (at most 2048, the finish kernels' and the no-allocation scratch's limit).
// Cache-resident data (op2h_streaming false for its 136 bytes per cell): g
// contiguous chunks of about three cells per thread, on XCD-ordered
// workgroups; streaming data: one cell per thread up to 2048 workgroups, a
// grid-stride sweep beyond.  chunk 0 = grid-stride.
struct UpdateShape {
  int g, chunk;
};
UpdateShape update_shape(int n, int bs) {
  if (n > 0 && !op2h_streaming((size_t)n * 136)) {
    const int g = std::max(1, std::min(grid_for(n, 3 * bs), 2048));
    return {g, (n + g - 1) / g};
  }
  return {std::max(1, std::min(grid_for(n, bs), 2048)), 0};
}

struct PendingDirect {
  bool upd = false, save = false, soa = false;
  op_set set = nullptr;
  hipStream_t st = nullptr;
  int n = 0, s = 0;
  // update (its partials handed to the graph's history kernel)
  int g = 0, bs = 0, chunk = 0;
  bool mixed = false;  // res and adt FP32 (res_f, adt_f), else FP64 (res, adt)
  const double *qold = nullptr;
  double *q = nullptr, *res = nullptr, *adt = nullptr;
  float *res_f = nullptr, *adt_f = nullptr;
  double *part_sum = nullptr, *part_max = nullptr;
  int *part_loc = nullptr;
  const int *oldidx = nullptr;
  // save_soln: src -> dst, n2 16-byte vectors when run on its own
  const double *save_src = nullptr;
  double *save_dst = nullptr;
  int save_n2 = 0, save_bs = 256;
};
PendingDirect g_direct;

void flush_direct(void *) {
  PendingDirect P = g_direct;
  g_direct = PendingDirect();
  if (P.upd && P.mixed)
    launch_k(false, nullptr, nullptr, P.soa ? k_update_mixed<true> : k_update_mixed<false>, P.g,
             P.bs, 0, P.st, P.qold, P.q, P.res_f, (const float *)P.adt_f, P.n, P.s, P.part_sum,
             P.part_max, P.part_loc, P.oldidx, P.chunk);
  else if (P.upd)
    launch_k(false, nullptr, nullptr, P.soa ? k_update<true> : k_update<false>, P.g, P.bs, 0,
             P.st, P.qold, P.q, P.res, (const double *)P.adt, P.n, P.s, P.part_sum,
             P.part_max, P.part_loc, P.oldidx, P.chunk);
  if (P.save)
    launch_k(false, nullptr, nullptr, k_save_soln, grid_for(P.save_n2, P.save_bs), P.save_bs, 0,
             P.st, (const double2 *)P.save_src, (double2 *)P.save_dst, P.save_n2);
}

// A pending fused direct chain ([update] [save_soln] adt_calc per cell).
template <class R>
using ChainKernel = void (*)(const double *, double *, R *, R *, double *, const double *,
                             const int *, const uint2 *, const int *, int, int, int, double *,
                             double *, int *, const int *, int, int);
struct PendingChain {
  ChainKernel<double> kern = nullptr;  // FP64 res and adt
  ChainKernel<float> kern_f = nullptr;  // mixed mode
  float *adt_f = nullptr;
  PendingDirect P;
  int g = 0, bs = 0, chunk = 0, n = 0, sq = 0, sx = 0;
  double *q = nullptr, *adt = nullptr;
  const double *x = nullptr;
  const int *pcell = nullptr;
  const uint2 *prow = nullptr;  // compact rows of pcell, or null
  const int *pbase = nullptr;
  hipStream_t st = nullptr;
};
PendingChain g_chain;

// res_dead: the update's zeros in res are not stored (the owner res_calc
// launched right after it sums from +0.0 over every cell of the set)
void launch_chain(bool res_dead) {
  const PendingChain C = g_chain;
  g_chain = PendingChain();
  const PendingDirect &P = C.P;
  if (C.kern_f)
    launch_k(false, nullptr, nullptr, C.kern_f, C.g, C.bs, 0, C.st, P.qold, C.q, P.res_f,
             C.adt_f, P.save_dst, C.x, C.pcell, C.prow, C.pbase, C.n, C.sq, C.sx, P.part_sum,
             P.part_max, P.part_loc, P.oldidx, C.chunk, (int)(res_dead && P.upd));
  else
    launch_k(false, nullptr, nullptr, C.kern, C.g, C.bs, 0, C.st, P.qold, C.q, P.res, C.adt,
             P.save_dst, C.x, C.pcell, C.prow, C.pbase, C.n, C.sq, C.sx, P.part_sum,
             P.part_max, P.part_loc, P.oldidx, C.chunk, (int)(res_dead && P.upd));
  KCHECK(hipGetLastError());
}

void flush_chain(void *) { launch_chain(false); }

// A pending owner-plan res_calc.
struct PendingRes {
  bool soa = false;
  hipStream_t st = nullptr;
  const op2h_owner_plan *plan = nullptr;
  OwnerArgs A{};
  const double *x = nullptr, *q = nullptr, *adt = nullptr;
  double *res = nullptr;
  int sx = 0, sq = 0, bs = 0;
  size_t lds = 0;
  op_map pecell = nullptr;
  bool zero = false;  // res holds the update's zeros: sums start from +0.0
};
PendingRes g_res;
PendingRes g_last_res;  // the latest owner-plan res_calc, pending or not

void launch_res_owner(const PendingRes &P) {
  if (P.plan->nblocks)
    launch_k(false, nullptr, nullptr,
             P.A.csr4 ? (P.soa ? k_res_owner<true, true> : k_res_owner<false, true>)
                      : (P.soa ? k_res_owner<true, false> : k_res_owner<false, false>),
             P.plan->nblocks, P.bs, P.lds, P.st, P.x, P.q, P.adt, P.res, P.A, P.sx, P.sq,
             (int)P.zero);
}

void flush_res(void *) {
  PendingRes P = g_res;
  g_res = PendingRes();
  launch_res_owner(P);
}

// The fused res_calc + bres_calc records for an owner plan and the boundary
// map (bedge -> cell), built on the host once per plan: each owned cell's
// interior increments, then its boundary edges (ascending; under transparent
// renumbering in the application's order).  Null when a cell has more than
// four increments in all, or the local ids do not fit 15 bits.
struct BoundaryRec {
  unsigned epoch = 0;  // op2h_plan_epoch() when built
  bool ok = false;
  BoundaryArgs B{};
};

// `build` false: only a record built before (inside a capture nothing may
// be allocated or uploaded).
const BoundaryArgs *boundary_args(const op2h_owner_plan *p, op_map pbecell, op_map pbedge,
                                  op_dat bound, hipStream_t st, bool build) {
  static std::map<std::pair<const void *, int>, BoundaryRec> cache;
  auto key = std::make_pair((const void *)p, pbecell->index);
  auto it = cache.find(key);
  if (it != cache.end() && it->second.epoch == op2h_plan_epoch()) {
    if (!it->second.ok) return nullptr;
    it->second.B.bound = (const int *)bound->data_d;
    it->second.B.pbedge = (const int2 *)pbedge->map_d;
    return &it->second.B;
  }
  if (!build) return nullptr;
  BoundaryRec &slot = cache[key];
  slot = BoundaryRec();
  slot.epoch = op2h_plan_epoch();
  if (!p->csr4 || pbecell->dim != 1 || (long long)p->max_elems * p->ninc >= 0x8000) return nullptr;
  const int nb = p->nblocks, no = p->n_owned;
  std::vector<int> oo(nb + 1), ow(std::max(no, 1)), eo(nb + 1), el(std::max(p->total, 1)),
      co(no + 1);
  op2h_owner_plan_download(p, oo.data(), ow.data(), eo.data(), el.data(), co.data(), nullptr);
  std::vector<unsigned short> csr(std::max(co[no], 1));
  op2h_owner_plan_download(p, nullptr, nullptr, nullptr, nullptr, nullptr, csr.data());
  // boundary edges per cell, in order
  const int nbe = pbecell->from->size, ncell = p->to_size;
  const int *key_be = op2h::old_index_host(pbecell->from);
  const int *becell = op2h::host_map(pbecell);  // in the device numbering
  std::vector<int> bo(ncell + 1, 0), order(nbe);
  for (int e = 0; e < nbe; ++e) bo[becell[e] + 1]++;
  for (int c = 0; c < ncell; ++c) bo[c + 1] += bo[c];
  for (int e = 0; e < nbe; ++e) order[e] = e;
  if (key_be)
    std::sort(order.begin(), order.end(), [&](int a, int b) { return key_be[a] < key_be[b]; });
  std::vector<int> fill(bo.begin(), bo.end() - 1), bl(std::max(nbe, 1));
  for (int e : order) bl[fill[becell[e]]++] = e;
  std::vector<ushort4> rec(std::max(no, 1));
  std::vector<int> offs(nb + 1, 0), edges;
  for (int b = 0; b < nb; ++b) {
    offs[b] = (int)edges.size();
    for (int i = oo[b]; i < oo[b + 1]; ++i) {
      unsigned short ent[4] = {0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF};
      int k = 0;
      for (int j = co[i]; j < co[i + 1]; ++j) {
        if (k == 4) return nullptr;
        ent[k++] = csr[j];
      }
      const int cell = ow[i];
      for (int j = bo[cell]; j < bo[cell + 1]; ++j) {
        const int local = (int)edges.size() - offs[b];
        if (k == 4 || local >= 0x7FFF) return nullptr;
        ent[k++] = (unsigned short)(0x8000 | local);
        edges.push_back(bl[j]);
      }
      rec[i] = make_ushort4(ent[0], ent[1], ent[2], ent[3]);
    }
  }
  offs[nb] = (int)edges.size();
  if (edges.size() != (size_t)nbe) return nullptr;  // a boundary edge of an unowned cell
  char name[96];
  snprintf(name, sizeof name, "res_bres.fused.%p.%d", (const void *)p, pbecell->index);
  const size_t b_rec = ((size_t)8 * std::max(no, 1) + 255) / 256 * 256;
  const size_t b_offs = ((size_t)4 * (nb + 1) + 255) / 256 * 256;
  int fresh = 0;
  char *buf = (char *)op2h_plan_user(name, b_rec + b_offs + 4 * std::max<size_t>(edges.size(), 1),
                                     &fresh);
  KCHECK(hipMemcpyAsync(buf, rec.data(), 8 * (size_t)no, hipMemcpyHostToDevice, st));
  KCHECK(hipMemcpyAsync(buf + b_rec, offs.data(), 4 * (size_t)(nb + 1), hipMemcpyHostToDevice, st));
  if (!edges.empty())
    KCHECK(hipMemcpyAsync(buf + b_rec + b_offs, edges.data(), 4 * edges.size(),
                          hipMemcpyHostToDevice, st));
  KCHECK(hipStreamSynchronize(st));
  BoundaryArgs &B = slot.B;
  B.rec = (const ushort4 *)buf;
  B.offs = (const int *)(buf + b_rec);
  B.edges = (const int *)(buf + b_rec + b_offs);
  B.pbedge = (const int2 *)pbedge->map_d;
  B.bound = (const int *)bound->data_d;
  slot.ok = true;
  return &B;
}

// Compact rows of a four-node map (cells -> nodes) for the fused direct
// chain: per cell its nodes as 16-bit offsets from the lowest node of its
// 64-cell group, 8 bytes instead of 16, where every group's nodes lie within
// 65535 of that base (a locally numbered mesh; else null).  `build` false:
// only a record built before (inside a capture nothing may be uploaded).
struct CompactRows {
  const uint2 *rec = nullptr;
  const int *base = nullptr;
};

const CompactRows *compact_rows(op_map m, hipStream_t st, bool build) {
  struct Slot {
    unsigned epoch = 0;
    bool ok = false;
    CompactRows R;
  };
  static std::map<int, Slot> cache;
  auto it = cache.find(m->index);
  if (it != cache.end() && it->second.epoch == op2h_plan_epoch())
    return it->second.ok ? &it->second.R : nullptr;
  if (!build) return nullptr;
  Slot &slot = cache[m->index];
  slot = Slot();
  slot.epoch = op2h_plan_epoch();
  const int n = m->from->size, ng = (n + 63) / 64;
  if (m->dim != 4 || n < 1) return nullptr;
  const int *hm = op2h::host_map(m);  // in the device numbering
  std::vector<uint2> rec(n);
  std::vector<int> base(ng);
  for (int g = 0; g < ng; ++g) {
    const int c0 = 64 * g, c1 = std::min(n, c0 + 64);
    int lo = INT_MAX, hi = INT_MIN;
    for (size_t i = 4 * (size_t)c0; i < 4 * (size_t)c1; ++i) {
      lo = std::min(lo, hm[i]);
      hi = std::max(hi, hm[i]);
    }
    if ((long long)hi - lo > 0xFFFF) return nullptr;
    base[g] = lo;
    for (int c = c0; c < c1; ++c) {
      const int *r = hm + 4 * (size_t)c;
      rec[c] = make_uint2((unsigned)(r[0] - lo) | ((unsigned)(r[1] - lo) << 16),
                          (unsigned)(r[2] - lo) | ((unsigned)(r[3] - lo) << 16));
    }
  }
  char name[64];
  snprintf(name, sizeof name, "compact_rows.%d", m->index);
  const size_t b_rec = (8 * (size_t)n + 255) / 256 * 256;
  int fresh = 0;
  char *buf = (char *)op2h_plan_user(name, b_rec + 4 * (size_t)ng, &fresh);
  KCHECK(hipMemcpyAsync(buf, rec.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st));
  KCHECK(hipMemcpyAsync(buf + b_rec, base.data(), 4 * (size_t)ng, hipMemcpyHostToDevice, st));
  KCHECK(hipStreamSynchronize(st));  // rec and base are locals
  slot.R.rec = (const uint2 *)buf;
  slot.R.base = (const int *)(buf + b_rec);
  slot.ok = true;
  return &slot.R;
}

}  // namespace

extern "C" void op_par_loop_airfoil_1_save_soln(char const *name, op_set set,
                                                op_arg arg0, op_arg arg1) {
  op_arg args[2] = {arg0, arg1};
  op2h_loop_ctx ctx;
  op2h_accept_deferred("direct_chain");
  op2h_loop_begin(&ctx, 0, name, set, 2, args, OP2H_PLAN_ATOMIC, 256);
  hipStream_t st = (hipStream_t)ctx.stream;
  const bool fuse = op2h_fusion_enabled() != 0;
  bool held = op2h_take_deferred("direct_chain", nullptr) != 0;
  if (held && !fuse) {  // launched first, in its place in the order
    flush_direct(nullptr);
    held = false;
  }
  if (set->size > 0) {
    int stride = 0;
    bool soa = loop_soa(args, 2, &stride);
    // identical layouts: the copy covers the whole device allocation
    long long ndouble = soa ? 4LL * set->stride : 4LL * set->size;
    int n2 = (int)(ndouble / 2);
    PendingDirect &P = g_direct;
    if (held && !(P.upd && !P.save && P.set == set && P.soa == soa &&
                  P.q == (double *)args[0].data_d)) {
      flush_direct(nullptr);  // not the update this copy follows
      held = false;
    }
    if (fuse) {  // held for the adt_calc that follows
      if (!held) {
        P = PendingDirect();
        P.set = set;
        P.soa = soa;
        P.st = st;
        P.n = set->size;
        P.s = soa ? set->stride : 0;
      }
      P.save = true;
      P.save_src = (const double *)args[0].data_d;
      P.save_dst = (double *)args[1].data_d;
      P.save_n2 = n2;
      P.save_bs = ctx.block_size;
      op2h_defer("direct_chain", flush_direct, nullptr);
      account(0, set, 2, args);
      op2h_loop_end(&ctx);
      return;
    }
    LoopTimer tm(&ctx);
    launch_k(tm.ext, tm.e0, tm.e1, k_save_soln, grid_for(n2, ctx.block_size), ctx.block_size,
             0, st, (const double2 *)args[0].data_d, (double2 *)args[1].data_d, n2);
    tm.done();
    KCHECK(hipGetLastError());
  } else if (held) {
    flush_direct(nullptr);
  }
  account(0, set, 2, args);
  op2h_loop_end(&ctx);
}

extern "C" void op_par_loop_airfoil_2_adt_calc(char const *name, op_set set,
                                               op_arg arg0, op_arg arg1,
                                               op_arg arg2, op_arg arg3,
                                               op_arg arg4, op_arg arg5) {
  op_arg args[6] = {arg0, arg1, arg2, arg3, arg4, arg5};
  op2h_loop_ctx ctx;
  op2h_accept_deferred("direct_chain");
  op2h_loop_begin(&ctx, 1, name, set, 6, args, OP2H_PLAN_ATOMIC, 256);
  hipStream_t st = (hipStream_t)ctx.stream;
  bool held = op2h_take_deferred("direct_chain", nullptr) != 0;
  if (held && !op2h_fusion_enabled()) {
    flush_direct(nullptr);
    held = false;
  }
  if (set->size > 0) {
    int dummy;
    bool soa = loop_soa(args, 6, &dummy);
    int sx = set_stride(args[0]), sq = set_stride(args[4]);
    const double *x = (const double *)args[0].data_d;
    const double *q = (const double *)args[4].data_d;
    double *adt = (double *)args[5].data_d;
    const int *pcell = args[0].map->map_d;
    int g = grid_for(set->size, ctx.block_size);
    if (held) {
      PendingDirect P = g_direct;
      const double *qsrc = P.upd ? P.q : P.save_src;
      const bool mixed = is_float(args[5]);
      if (P.set == set && P.soa == soa && (soa ? sq == P.s : true) && qsrc == q &&
          (!P.upd || (P.mixed == mixed &&
                      (mixed ? (void *)P.adt_f == args[5].data_d : P.adt == adt)))) {
        g_direct = PendingDirect();
        // [update] [save_soln] adt_calc per cell: the update's launch shape
        // (its partials), else adt_calc's
        const int gg = P.upd ? P.g : g, bb = P.upd ? P.bs : ctx.block_size;
        const int chunk = P.upd ? P.chunk : 0;
        PendingChain &C = g_chain;
        C = PendingChain();
        if (mixed) {
          C.kern_f = P.upd ? (P.save ? (soa ? k_direct_chain<float, true, true, true> : k_direct_chain<float, false, true, true>)
                                     : (soa ? k_direct_chain<float, true, true, false> : k_direct_chain<float, false, true, false>))
                           : (soa ? k_direct_chain<float, true, false, true> : k_direct_chain<float, false, false, true>);
          C.adt_f = (float *)args[5].data_d;
        } else {
          C.kern = P.upd ? (P.save ? (soa ? k_direct_chain<double, true, true, true> : k_direct_chain<double, false, true, true>)
                                   : (soa ? k_direct_chain<double, true, true, false> : k_direct_chain<double, false, true, false>))
                         : (soa ? k_direct_chain<double, true, false, true> : k_direct_chain<double, false, false, true>);
        }
        C.P = P;
        C.g = gg;
        C.bs = bb;
        C.chunk = chunk;
        C.q = P.q ? P.q : (double *)q;
        C.adt = adt;
        C.x = x;
        C.pcell = pcell;
        if (const CompactRows *cr = compact_rows(args[0].map, st, false)) {
          C.prow = cr->rec;
          C.pbase = cr->base;
        }
        C.n = set->size;
        C.sq = sq;
        C.sx = sx;
        C.st = st;
        // held for the res_calc that follows, which may take the update's
        // zeros in res as known (launched as it is by anything else)
        op2h_defer("chain", flush_chain, nullptr);
        account(1, set, 6, args);
        op2h_loop_end(&ctx);
        return;
      }
      flush_direct(nullptr);
    }
    // the chains fused inside a captured step read the compact rows, built
    // here (outside capture) at the first eager call
    if (!ctx.capturing) (void)compact_rows(args[0].map, st, true);
    LoopTimer tm(&ctx);
    if (is_float(args[5])) {
      float *adtf = (float *)args[5].data_d;
      launch_k(tm.ext, tm.e0, tm.e1, soa ? k_adt_calc_mixed<true> : k_adt_calc_mixed<false>, g,
               ctx.block_size, 0, st, x, q, adtf, pcell, set->size, sx, sq);
    } else {
      // streams: pcell rows, q and adt
      const bool nt = op2h_streaming((size_t)set->size * 56) != 0;
      launch_k(tm.ext, tm.e0, tm.e1,
               nt ? (soa ? k_adt_calc<true, true> : k_adt_calc<false, true>)
                  : (soa ? k_adt_calc<true, false> : k_adt_calc<false, false>),
               g, ctx.block_size, 0, st, x, q, adt, pcell, set->size, sx, sq);
    }
    tm.done();
    KCHECK(hipGetLastError());
  } else if (held) {
    flush_direct(nullptr);
  }
  account(1, set, 6, args);
  op2h_loop_end(&ctx);
}

extern "C" void op_par_loop_airfoil_3_res_calc(char const *name, op_set set,
                                               op_arg arg0, op_arg arg1,
                                               op_arg arg2, op_arg arg3,
                                               op_arg arg4, op_arg arg5,
                                               op_arg arg6, op_arg arg7) {
  op_arg args[8] = {arg0, arg1, arg2, arg3, arg4, arg5, arg6, arg7};
  op2h_loop_ctx ctx;
  op2h_accept_deferred("chain");
  op2h_loop_begin(&ctx, 2, name, set, 8, args, OP2H_PLAN_STAGED, 256);
  hipStream_t st = (hipStream_t)ctx.stream;
  // a fused direct chain held for this loop: launched here, before it, with
  // or without the update's zeros in res (only an owner plan covering every
  // cell can take them as known)
  bool chain = op2h_take_deferred("chain", nullptr) != 0;
  if (chain && (!op2h_fusion_enabled() || set->size == 0 || ctx.plan != OP2H_PLAN_OWNER)) {
    flush_chain(nullptr);
    chain = false;
  }
  if (set->size > 0) {
    int dummy;
    bool soa = loop_soa(args, 8, &dummy);
    int sx = set_stride(args[0]), sq = set_stride(args[2]);
    const double *x = (const double *)args[0].data_d;
    const double *q = (const double *)args[2].data_d;
    const double *adt = (const double *)args[4].data_d;
    double *res = (double *)args[6].data_d;
    const int2 *pedge = (const int2 *)args[0].map->map_d;
    const int2 *pecell = (const int2 *)args[2].map->map_d;
    const int bs = ctx.block_size;
    const bool mixed = is_float(args[6]);
    if (mixed != is_float(args[4])) {
      printf("op2hip: res_calc needs adt and res both FP32 (mixed mode) or both FP64\n");
      exit(-1);
    }
    if (mixed) {
      // mixed precision: atomic plan, else the owner-compute plan
      const float *adtf = (const float *)args[4].data_d;
      float *resf = (float *)args[6].data_d;
      const op2h_owner_plan *owner_p = nullptr;
      OwnerArgs OA{};
      bool zero = false;  // as in FP64 below: res dead after the chain's update
      if (ctx.plan == OP2H_PLAN_OWNER) {  // built before the timed region
        ctx.part_size = bs;
        owner_p = op2h_plan_owner_local(&ctx);
        OA = owner_args_res(owner_p, args, st);
        zero = chain && g_chain.P.upd && g_chain.P.mixed && g_chain.P.res_f == resf &&
               owner_p->n_owned == args[6].dat->set->size;
      }
      if (chain) launch_chain(zero);
      chain = false;
      op2h_kernel_timer_start(&ctx);
      if (ctx.plan == OP2H_PLAN_ATOMIC) {
        int g = grid_for(set->size, bs);
        if (soa) k_res_atomic_mixed<true><<<g, bs, 0, st>>>(x, q, adtf, resf, pedge, pecell, set->size, sx, sq);
        else k_res_atomic_mixed<false><<<g, bs, 0, st>>>(x, q, adtf, resf, pedge, pecell, set->size, sx, sq);
      } else if (ctx.plan == OP2H_PLAN_OWNER) {
        const int nblocks = owner_p->nblocks;
        const size_t lds = sizeof(float) * 4 * (size_t)OA.max_e;
        if (lds > 160 * 1024) {
          printf("op2hip: res_calc owner plan needs %zu bytes of LDS; use a smaller block\n", lds);
          exit(-1);
        }
        if (nblocks) {
          auto kern = OA.csr4 ? (soa ? k_res_owner_mixed<true, true> : k_res_owner_mixed<false, true>)
                              : (soa ? k_res_owner_mixed<true, false> : k_res_owner_mixed<false, false>);
          launch_k(false, nullptr, nullptr, kern, nblocks, bs, lds, st, x, q, adtf, resf, OA,
                   sx, sq, (int)zero);
        }
      } else {
        const op2h_gather_plan *p = op2h_plan_gather(args[6].map);
        if (p->max_count > 4) {
          printf("op2hip: mixed-precision res_calc supports at most 4 edges per cell\n");
          exit(-1);
        }
        char key[96];
        snprintf(key, sizeof key, "res_calc.gather4.%d.%d", args[0].map->index,
                 args[2].map->index);
        int fresh = 0;
        int4 *rec = (int4 *)op2h_plan_user(key, sizeof(int4) * 4 * (size_t)p->ntargets, &fresh);
        if (fresh)
          k_pack_res_gather<4><<<grid_for(4 * (long long)p->ntargets, 256), 256, 0, st>>>(
              pedge, pecell, p->offsets, p->entries, p->ntargets, rec);
        int g = grid_for(4 * (long long)p->ntargets, bs);
        if (soa) k_res_gather_packed_mixed<true, 4><<<g, bs, 0, st>>>(x, q, adtf, resf, rec, p->targets, p->ntargets, sx, sq);
        else k_res_gather_packed_mixed<false, 4><<<g, bs, 0, st>>>(x, q, adtf, resf, rec, p->targets, p->ntargets, sx, sq);
      }
      op2h_kernel_timer_stop(&ctx);
      KCHECK(hipGetLastError());
      account(2, set, 8, args);
      op2h_loop_end(&ctx);
      return;
    }
    switch (ctx.plan) {
      case OP2H_PLAN_ATOMIC: {
        op2h_kernel_timer_start(&ctx);
        int g = grid_for(set->size, bs);
        if (soa) k_res_atomic<true><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, set->size, sx, sq);
        else k_res_atomic<false><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, set->size, sx, sq);
        op2h_kernel_timer_stop(&ctx);
        break;
      }
      case OP2H_PLAN_COLOR: {
        const op2h_color_plan *p = op2h_plan_color(&ctx);
        op2h_kernel_timer_start(&ctx);
        for (int c = 0; c < p->ncolors; ++c) {
          int s0 = p->col_offsets_h[c], s1 = p->col_offsets_h[c + 1];
          int g = grid_for(s1 - s0, bs);
          if (!g) continue;
          if (soa) k_res_color<true><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, p->col_reord, s0, s1, sx, sq);
          else k_res_color<false><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, p->col_reord, s0, s1, sx, sq);
        }
        op2h_kernel_timer_stop(&ctx);
        break;
      }
      case OP2H_PLAN_GATHER: {
        const op2h_gather_plan *p = op2h_plan_gather(args[6].map);
        LoopTimer tm(&ctx);
        if (p->max_count <= 4) {
          // packed per-(cell, edge) records, built once per plan
          char key[96];
          snprintf(key, sizeof key, "res_calc.gather4.%d.%d", args[0].map->index,
                   args[2].map->index);
          int fresh = 0;
          int4 *rec = (int4 *)op2h_plan_user(key, sizeof(int4) * 4 * (size_t)p->ntargets, &fresh);
          if (fresh)
            k_pack_res_gather<4><<<grid_for(4 * (long long)p->ntargets, 256), 256, 0, st>>>(
                pedge, pecell, p->offsets, p->entries, p->ntargets, rec);
          int g = grid_for(4 * (long long)p->ntargets, bs);
          launch_k(tm.ext, tm.e0, tm.e1,
                   soa ? k_res_gather_packed<true, 4> : k_res_gather_packed<false, 4>, g, bs, 0,
                   st, x, q, adt, res, rec, p->targets, p->ntargets, sx, sq);
        } else if (p->max_count <= 8) {
          if (tm.ext) KCHECK(hipEventRecord((hipEvent_t)tm.e0, st));
          int g = grid_for(8 * (long long)p->ntargets, bs);
          if (soa) k_res_gather_lanes<true, 8><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
          else k_res_gather_lanes<false, 8><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
        } else {
          if (tm.ext) KCHECK(hipEventRecord((hipEvent_t)tm.e0, st));
          int g = grid_for(p->ntargets, bs);
          if (soa) k_res_gather<true><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
          else k_res_gather<false><<<g, bs, 0, st>>>(x, q, adt, res, pedge, pecell, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
        }
        if (tm.ext && p->max_count > 4) KCHECK(hipEventRecord((hipEvent_t)tm.e1, st));
        tm.done();
        break;
      }
      case OP2H_PLAN_OWNER: {
        ctx.part_size = bs;
        const op2h_owner_plan *p = op2h_plan_owner_local(&ctx);
        OwnerArgs A = owner_args_res(p, args, st);
        const size_t lds = sizeof(double) * 4 * (size_t)A.max_e;
        if (lds > 160 * 1024) {
          printf("op2hip: res_calc owner plan needs %zu bytes of LDS; use a smaller block\n", lds);
          exit(-1);
        }
        // res is dead between the chain's update (which sets every cell's
        // res to +0.0) and this loop when the plan owns every cell: the
        // zeros are neither stored nor read
        const bool zero = chain && g_chain.P.upd && g_chain.P.res == res &&
                          p->n_owned == args[6].dat->set->size;
        if (chain) launch_chain(zero);
        chain = false;
        PendingRes R;
        R.soa = soa;
        R.st = st;
        R.plan = p;
        R.A = A;
        R.x = x;
        R.q = q;
        R.adt = adt;
        R.res = res;
        R.sx = sx;
        R.sq = sq;
        R.bs = bs;
        R.lds = lds;
        R.pecell = args[6].map;
        R.zero = zero;
        g_last_res = R;
        if (op2h_fusion_enabled()) {  // held for the bres_calc that follows
          g_res = R;
          op2h_defer("res_owner", flush_res, nullptr);
          break;
        }
        LoopTimer tm(&ctx);
        if (p->nblocks)
          launch_k(tm.ext, tm.e0, tm.e1,
                   A.csr4 ? (soa ? k_res_owner<true, true> : k_res_owner<false, true>)
                          : (soa ? k_res_owner<true, false> : k_res_owner<false, false>),
                   p->nblocks, bs, lds, st, x, q, adt, res, A, sx, sq, (int)zero);
        tm.done();
        break;
      }
      default: {  // STAGED
        ctx.part_size = bs;
        const op2h_staged_plan *p = op2h_plan_staged(&ctx);
        int sn = p->map_of_arg[0], sc = p->map_of_arg[2];
        StagedArgs A;
        A.blkmap = p->blkmap;
        A.offset = p->offset;
        A.nelems = p->nelems;
        A.nthrcol = p->nthrcol;
        A.thrcol = p->thrcol;
        A.ind_offs_n = p->ind_offs[sn];
        A.ind_sizes_n = p->ind_sizes[sn];
        A.ind_map_n = p->ind_map[sn];
        A.ind_offs_c = p->ind_offs[sc];
        A.ind_sizes_c = p->ind_sizes[sc];
        A.ind_map_c = p->ind_map[sc];
        A.loc_n0 = p->loc_map[p->loc_of_arg[0]];
        A.loc_n1 = p->loc_map[p->loc_of_arg[1]];
        A.loc_c0 = p->loc_map[p->loc_of_arg[2]];
        A.loc_c1 = p->loc_map[p->loc_of_arg[3]];
        A.max_n = p->max_ind_h[sn];
        A.max_c = p->max_ind_h[sc];
        size_t lds = sizeof(double) * (2 * (size_t)A.max_n + 9 * (size_t)A.max_c);
        if (lds > 160 * 1024) {
          printf("op2hip: res_calc staged plan needs %zu bytes of LDS\n", lds);
          exit(-1);
        }
        op2h_kernel_timer_start(&ctx);
        for (int c = 0; c < p->ncolors; ++c) {
          int nb = p->ncolblk_h[c];
          if (!nb) continue;
          if (soa) k_res_staged<true><<<nb, bs, lds, st>>>(x, q, adt, res, A, p->colstart_h[c], sx, sq);
          else k_res_staged<false><<<nb, bs, lds, st>>>(x, q, adt, res, A, p->colstart_h[c], sx, sq);
        }
        op2h_kernel_timer_stop(&ctx);
        break;
      }
    }
    KCHECK(hipGetLastError());
  }
  if (chain) flush_chain(nullptr);  // (every path above has launched it)
  account(2, set, 8, args);
  op2h_loop_end(&ctx);
}

extern "C" void op_par_loop_airfoil_4_bres_calc(char const *name, op_set set,
                                                op_arg arg0, op_arg arg1,
                                                op_arg arg2, op_arg arg3,
                                                op_arg arg4, op_arg arg5) {
  op_arg args[6] = {arg0, arg1, arg2, arg3, arg4, arg5};
  op2h_loop_ctx ctx;
  op2h_accept_deferred("res_owner");
  op2h_loop_begin(&ctx, 3, name, set, 6, args, OP2H_PLAN_STAGED, 256);
  hipStream_t st = (hipStream_t)ctx.stream;
  bool held = op2h_take_deferred("res_owner", nullptr) != 0;
  if (held && !op2h_fusion_enabled()) {
    flush_res(nullptr);
    held = false;
  }
  if (set->size > 0) {
    int dummy;
    bool soa = loop_soa(args, 6, &dummy);
    int sx = set_stride(args[0]), sq = set_stride(args[2]);
    const double *x = (const double *)args[0].data_d;
    const double *q = (const double *)args[2].data_d;
    const double *adt = (const double *)args[3].data_d;
    double *res = (double *)args[4].data_d;
    const int *bound = (const int *)args[5].data_d;
    const int2 *pbedge = (const int2 *)args[0].map->map_d;
    const int *pbecell = args[2].map->map_d;
    const int bs = ctx.block_size;
    // the res_calc this loop follows, on the same dats: the two as one owner
    // launch (inside a capture), its boundary records built beforehand
    const PendingRes &L = held ? g_res : g_last_res;
    const bool same = L.plan && L.x == x && L.q == q && L.adt == adt && L.res == res &&
                      L.soa == soa && L.sx == sx && L.sq == sq && !is_float(args[4]) &&
                      args[2].map->to == L.pecell->to && args[4].map == args[2].map;
    const BoundaryArgs *B = same ? boundary_args(L.plan, args[2].map, args[0].map, args[5].dat,
                                                 st, !ctx.capturing)
                                 : nullptr;
    if (held) {
      if (B) {
        PendingRes P = g_res;
        g_res = PendingRes();
        if (P.plan->nblocks)
          launch_k(false, nullptr, nullptr, soa ? k_res_bres_owner<true> : k_res_bres_owner<false>,
                   P.plan->nblocks, P.bs, P.lds, st, x, q, adt, res, P.A, *B, sx, sq,
                   (int)P.zero);
        KCHECK(hipGetLastError());
        account(3, set, 6, args);
        op2h_loop_end(&ctx);
        return;
      }
      flush_res(nullptr);
    }
    if (is_float(args[4])) {
      const float *adtf = (const float *)args[3].data_d;
      float *resf = (float *)args[4].data_d;
      op2h_kernel_timer_start(&ctx);
      if (ctx.plan == OP2H_PLAN_ATOMIC) {
        int g = grid_for(set->size, bs);
        if (soa) k_bres_atomic_mixed<true><<<g, bs, 0, st>>>(x, q, adtf, resf, pbedge, pbecell, bound, set->size, sx, sq);
        else k_bres_atomic_mixed<false><<<g, bs, 0, st>>>(x, q, adtf, resf, pbedge, pbecell, bound, set->size, sx, sq);
      } else {
        const op2h_gather_plan *p = op2h_plan_gather(args[4].map);
        int g = grid_for(p->ntargets, bs);
        if (soa) k_bres_gather_mixed<true><<<g, bs, 0, st>>>(x, q, adtf, resf, pbedge, bound, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
        else k_bres_gather_mixed<false><<<g, bs, 0, st>>>(x, q, adtf, resf, pbedge, bound, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
      }
      op2h_kernel_timer_stop(&ctx);
      KCHECK(hipGetLastError());
      account(3, set, 6, args);
      op2h_loop_end(&ctx);
      return;
    }
    switch (ctx.plan) {
      case OP2H_PLAN_ATOMIC: {
        LoopTimer tm(&ctx);
        int g = grid_for(set->size, bs);
        launch_k(tm.ext, tm.e0, tm.e1, soa ? k_bres_atomic<true> : k_bres_atomic<false>, g, bs,
                 0, st, x, q, adt, res, pbedge, pbecell, bound, set->size, sx, sq);
        tm.done();
        break;
      }
      case OP2H_PLAN_COLOR: {
        const op2h_color_plan *p = op2h_plan_color(&ctx);
        op2h_kernel_timer_start(&ctx);
        for (int c = 0; c < p->ncolors; ++c) {
          int s0 = p->col_offsets_h[c], s1 = p->col_offsets_h[c + 1];
          int g = grid_for(s1 - s0, bs);
          if (!g) continue;
          if (soa) k_bres_color<true><<<g, bs, 0, st>>>(x, q, adt, res, pbedge, pbecell, bound, p->col_reord, s0, s1, sx, sq);
          else k_bres_color<false><<<g, bs, 0, st>>>(x, q, adt, res, pbedge, pbecell, bound, p->col_reord, s0, s1, sx, sq);
        }
        op2h_kernel_timer_stop(&ctx);
        break;
      }
      // One increment per boundary edge: an owner partition would evaluate
      // each edge once in its cell's block, which is what the gather plan does.
      case OP2H_PLAN_OWNER:
      case OP2H_PLAN_GATHER: {
        const op2h_gather_plan *p = op2h_plan_gather(args[4].map);
        op2h_kernel_timer_start(&ctx);
        int g = grid_for(p->ntargets, bs);
        if (soa) k_bres_gather<true><<<g, bs, 0, st>>>(x, q, adt, res, pbedge, bound, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
        else k_bres_gather<false><<<g, bs, 0, st>>>(x, q, adt, res, pbedge, bound, p->offsets, p->entries, p->targets, p->ntargets, sx, sq);
        op2h_kernel_timer_stop(&ctx);
        break;
      }
      default: {  // STAGED
        ctx.part_size = bs;
        const op2h_staged_plan *p = op2h_plan_staged(&ctx);
        int sn = p->map_of_arg[0], sc = p->map_of_arg[2];
        BStagedArgs A;
        A.blkmap = p->blkmap;
        A.offset = p->offset;
        A.nelems = p->nelems;
        A.nthrcol = p->nthrcol;
        A.thrcol = p->thrcol;
        A.ind_offs_n = p->ind_offs[sn];
        A.ind_sizes_n = p->ind_sizes[sn];
        A.ind_map_n = p->ind_map[sn];
        A.ind_offs_c = p->ind_offs[sc];
        A.ind_sizes_c = p->ind_sizes[sc];
        A.ind_map_c = p->ind_map[sc];
        A.loc_n0 = p->loc_map[p->loc_of_arg[0]];
        A.loc_n1 = p->loc_map[p->loc_of_arg[1]];
        A.loc_c0 = p->loc_map[p->loc_of_arg[2]];
        A.max_n = p->max_ind_h[sn];
        A.max_c = p->max_ind_h[sc];
        size_t lds = sizeof(double) * (2 * (size_t)A.max_n + 9 * (size_t)A.max_c);
        if (lds > 160 * 1024) {
          printf("op2hip: bres_calc staged plan needs %zu bytes of LDS\n", lds);
          exit(-1);
        }
        op2h_kernel_timer_start(&ctx);
        for (int c = 0; c < p->ncolors; ++c) {
          int nb = p->ncolblk_h[c];
          if (!nb) continue;
          if (soa) k_bres_staged<true><<<nb, bs, lds, st>>>(x, q, adt, res, bound, A, p->colstart_h[c], sx, sq);
          else k_bres_staged<false><<<nb, bs, lds, st>>>(x, q, adt, res, bound, A, p->colstart_h[c], sx, sq);
        }
        op2h_kernel_timer_stop(&ctx);
        break;
      }
    }
    KCHECK(hipGetLastError());
  } else if (held) {
    flush_res(nullptr);
  }
  account(3, set, 6, args);
  op2h_loop_end(&ctx);
}

extern "C" void op_par_loop_airfoil_5_update(char const *name, op_set set,
                                             op_arg arg0, op_arg arg1,
                                             op_arg arg2, op_arg arg3,
                                             op_arg arg4, op_arg arg5,
                                             op_arg arg6, op_arg arg7) {
  op_arg args[8] = {arg0, arg1, arg2, arg3, arg4, arg5, arg6, arg7};
  op2h_loop_ctx ctx;
  op2h_loop_begin(&ctx, 4, name, set, 8, args, OP2H_PLAN_ATOMIC, 256);
  hipStream_t st = (hipStream_t)ctx.stream;
  int dummy;
  bool soa = loop_soa(args, 8, &dummy);
  int s = set_stride(args[0]);
  const int bs = ctx.block_size;
  const UpdateShape shape = update_shape(set->size, bs);
  const int g = shape.g;
  // inside a captured graph the partials go to the runtime, whose history
  // kernel finishes both reductions of every update of the step in one pass
  // (same tree as k_update_finish); otherwise finish them here
  double *part_sum = nullptr, *part_max = nullptr, *slot_rms = nullptr, *slot_max = nullptr;
  int *part_loc = nullptr, *unused = nullptr;
  const bool deferred = op2h_gbl_partials(&ctx, 4, -1, g, &part_sum, &unused) != 0;
  if (deferred && !op2h_gbl_partials(&ctx, 5, 7, g, &part_max, &part_loc)) {
    printf("op2hip: update: maxerr/errloc cannot be deferred with rms\n");
    exit(-1);
  }
  if (!deferred) {
    slot_rms = (double *)op2h_gbl_slot(&ctx, 4, -1);
    slot_max = (double *)op2h_gbl_slot(&ctx, 5, 7);
    char *partials = (char *)op2h_reduct_scratch((size_t)g * 24);
    part_sum = (double *)partials;
    part_max = part_sum + g;
    part_loc = (int *)(part_max + g);
  }
  if (is_float(args[2]) != is_float(args[3])) {
    printf("op2hip: update needs res and adt both FP32 (mixed mode) or both FP64\n");
    exit(-1);
  }
  if (deferred && op2h_fusion_enabled() && set->size > 0) {
    // held for the save_soln / adt_calc that follow (its partials are
    // finished by the graph's history kernel either way)
    PendingDirect &P = g_direct;
    P = PendingDirect();
    P.upd = true;
    P.mixed = is_float(args[2]);
    P.set = set;
    P.soa = soa;
    P.st = st;
    P.n = set->size;
    P.s = s;
    P.g = g;
    P.bs = bs;
    P.chunk = shape.chunk;
    P.qold = (const double *)args[0].data_d;
    P.q = (double *)args[1].data_d;
    if (P.mixed) {
      P.res_f = (float *)args[2].data_d;
      P.adt_f = (float *)args[3].data_d;
    } else {
      P.res = (double *)args[2].data_d;
      P.adt = (double *)args[3].data_d;
    }
    P.part_sum = part_sum;
    P.part_max = part_max;
    P.part_loc = part_loc;
    P.oldidx = op2h_set_old_index(set);
    op2h_defer("direct_chain", flush_direct, nullptr);
    account(4, set, 8, args);
    op2h_loop_end(&ctx);
    return;
  }
  LoopTimer tm(&ctx);
  if (is_float(args[2])) {
    launch_k(tm.ext, tm.e0, nullptr, soa ? k_update_mixed<true> : k_update_mixed<false>, g, bs,
             0, st, (const double *)args[0].data_d, (double *)args[1].data_d,
             (float *)args[2].data_d, (const float *)args[3].data_d, set->size, s, part_sum,
             part_max, part_loc, op2h_set_old_index(set), shape.chunk);
  } else {
    launch_k(tm.ext, tm.e0, nullptr, soa ? k_update<true> : k_update<false>, g, bs, 0, st,
             (const double *)args[0].data_d, (double *)args[1].data_d,
             (double *)args[2].data_d, (const double *)args[3].data_d, set->size, s,
             part_sum, part_max, part_loc, op2h_set_old_index(set), shape.chunk);
  }
  if (!deferred)
    launch_k(tm.ext, nullptr, tm.e1, k_update_finish, 1, kFinishThreads, 0, st, part_sum,
             part_max, part_loc, g, slot_rms, slot_max);
  tm.done();
  KCHECK(hipGetLastError());
  account(4, set, 8, args);
  op2h_loop_end(&ctx);
}

// ---------------------------------------------------------------------------
// The update loop on caller-owned device arrays without allocating: the same
// kernels as op_par_loop_airfoil_5_update (same workgroup count for a block
// size, so the same reduction tree and the same bits), with the per-workgroup
// partials in a file-scope device array instead of runtime-allocated memory.
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid it).  Every call on one device
// shares that device's array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
namespace {
constexpr int kNoAllocMaxParts = 2048;  // the stub's workgroup cap
constexpr size_t kNoAllocScratchBytes = (size_t)kNoAllocMaxParts * (8 + 8 + 4);  // 40 KiB
constexpr int kNoAllocMaxDevices = 64;
void *g_scratch_addr[kNoAllocMaxDevices];  // device address of `scratch`, per device
}  // namespace

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

extern "C" int op2h_update_noalloc(int ncell, int soa_stride, int block, const double *qold,
                                   double *q, double *res, const double *adt,
                                   op2h_update_result *result, void *stream) {
  if (ncell < 0 || block < 64 || block > 1024 || block % 64 != 0 || !result ||
      (ncell > 0 && (!qold || !q || !res || !adt)) ||
      (soa_stride != 0 && soa_stride < ncell)) {
    printf("op2h_update_noalloc: bad arguments (ncell %d, soa_stride %d, block %d)\n", ncell,
           soa_stride, block);
    return -1;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
    printf("op2h_update_noalloc: no usable current device\n");
    return -2;
  }
  if (!g_scratch_addr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      printf("op2h_update_noalloc: hipGetSymbolAddress failed: %s\n", hipGetErrorString(e));
      return -2;
    }
    g_scratch_addr[dev] = ptr;
  }
  const UpdateShape shape = update_shape(ncell, block);  // the stub's
  const int g = shape.g;
  if ((size_t)g * (8 + 8 + 4) > kNoAllocScratchBytes) {
    printf("op2h_update_noalloc: %d partials exceed the %zu-byte scratch\n", g,
           kNoAllocScratchBytes);
    return -3;
  }
  double *part_sum = (double *)g_scratch_addr[dev];
  double *part_max = part_sum + g;
  int *part_loc = (int *)(part_max + g);
  hipStream_t st = (hipStream_t)stream;
  // every partial the finish kernel reads is written by this call's k_update
  if (soa_stride)
    k_update<true><<<g, block, 0, st>>>(qold, q, res, adt, ncell, soa_stride, part_sum,
                                        part_max, part_loc, nullptr, shape.chunk);
  else
    k_update<false><<<g, block, 0, st>>>(qold, q, res, adt, ncell, 0, part_sum, part_max,
                                         part_loc, nullptr, shape.chunk);
  k_update_finish<<<1, kFinishThreads, 0, st>>>(part_sum, part_max, part_loc, g, &result->rms,
                                               &result->maxerr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    printf("op2h_update_noalloc: launch failed: %s\n", hipGetErrorString(e));
    return -4;
  }
  return 0;
}
