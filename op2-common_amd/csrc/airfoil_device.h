// Device-side pieces of the airfoil application shared by the single-mesh
// loops (airfoil_kernels.hip) and the batched ensemble (ensemble.hip): the
// application's constants, its user kernels and the dat access helpers.
// Constants have internal linkage: each translation unit owns its copies and
// sets them itself.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

// ---------------------------------------------------------------------------
// application constants (op_decl_const, airfoil.cpp:204-210)
// ---------------------------------------------------------------------------
static __constant__ double gam_d;
static __constant__ double gm1_d;
static __constant__ double cfl_d;
static __constant__ double eps_d;
static __constant__ double mach_d;
static __constant__ double alpha_d;
static __constant__ double qinf_d[4];
// FP32 copies for the mixed-precision flux kernels
static __constant__ float gam_f;
static __constant__ float gm1_f;
static __constant__ float cfl_f;
static __constant__ float eps_f;
static __constant__ float qinf_f[4];

// ---------------------------------------------------------------------------
// user kernels (apps/c/airfoil/airfoil_plain/dp/*.h)
// ---------------------------------------------------------------------------
namespace op2_k {

__device__ __forceinline__ void adt_calc(const double *x1, const double *x2,
                                         const double *x3, const double *x4,
                                         const double *q, double *adt) {
  double dx, dy, ri, u, v, c;
  ri = 1.0 / q[0];
  u = ri * q[1];
  v = ri * q[2];
  c = sqrt(gam_d * gm1_d * (ri * q[3] - 0.5 * (u * u + v * v)));
  dx = x2[0] - x1[0];
  dy = x2[1] - x1[1];
  *adt = fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  dx = x3[0] - x2[0];
  dy = x3[1] - x2[1];
  *adt += fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  dx = x4[0] - x3[0];
  dy = x4[1] - x3[1];
  *adt += fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  dx = x1[0] - x4[0];
  dy = x1[1] - x4[1];
  *adt += fabs(u * dy - v * dx) + c * sqrt(dx * dx + dy * dy);
  *adt = (*adt) / cfl_d;
}

__device__ __forceinline__ void res_calc(const double *x1, const double *x2,
                                         const double *q1, const double *q2,
                                         const double *adt1, const double *adt2,
                                         double *res1, double *res2) {
  double dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0 / q1[0];
  p1 = gm1_d * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0 / q2[0];
  p2 = gm1_d * (q2[3] - 0.5 * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  vol2 = ri * (q2[1] * dy - q2[2] * dx);
  mu = 0.5 * ((*adt1) + (*adt2)) * eps_d;
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

// The increment f of res_calc (res1 += f, res2 -= f), with res_calc's
// operations in res_calc's order: f is bit-identical to what res_calc adds.
__device__ __forceinline__ void res_flux(const double *x1, const double *x2,
                                         const double *q1, const double *q2,
                                         double adt1, double adt2, double *f) {
  double dx, dy, mu, ri, p1, vol1, p2, vol2;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0 / q1[0];
  p1 = gm1_d * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0 / q2[0];
  p2 = gm1_d * (q2[3] - 0.5 * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  vol2 = ri * (q2[1] * dy - q2[2] * dx);
  mu = 0.5 * (adt1 + adt2) * eps_d;
  f[0] = 0.5 * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  f[1] = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) + mu * (q1[1] - q2[1]);
  f[2] = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) + mu * (q1[2] - q2[2]);
  f[3] = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
}

// bres_calc with the free stream qi (qinf of the program, or a member's own)
__device__ __forceinline__ void bres_calc_q(const double *x1, const double *x2,
                                            const double *q1, const double *adt1,
                                            double *res1, const int *bound,
                                            const double *qi) {
  double dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0 / q1[0];
  p1 = gm1_d * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  if (*bound == 1) {
    res1[1] += +p1 * dy;
    res1[2] += -p1 * dx;
  } else {
    vol1 = ri * (q1[1] * dy - q1[2] * dx);
    ri = 1.0 / qi[0];
    p2 = gm1_d * (qi[3] - 0.5 * ri * (qi[1] * qi[1] + qi[2] * qi[2]));
    vol2 = ri * (qi[1] * dy - qi[2] * dx);
    mu = (*adt1) * eps_d;
    f = 0.5 * (vol1 * q1[0] + vol2 * qi[0]) + mu * (q1[0] - qi[0]);
    res1[0] += f;
    f = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * qi[1] + p2 * dy) + mu * (q1[1] - qi[1]);
    res1[1] += f;
    f = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * qi[2] - p2 * dx) + mu * (q1[2] - qi[2]);
    res1[2] += f;
    f = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (qi[3] + p2)) + mu * (q1[3] - qi[3]);
    res1[3] += f;
  }
}

// bres_calc.h with the program's qinf (kept apart from bres_calc_q so that
// qinf stays in the constant bank: scalar loads, no vector registers)
__device__ __forceinline__ void bres_calc(const double *x1, const double *x2,
                                          const double *q1, const double *adt1,
                                          double *res1, const int *bound) {
  double dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0 / q1[0];
  p1 = gm1_d * (q1[3] - 0.5 * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  if (*bound == 1) {
    res1[1] += +p1 * dy;
    res1[2] += -p1 * dx;
  } else {
    vol1 = ri * (q1[1] * dy - q1[2] * dx);
    ri = 1.0 / qinf_d[0];
    p2 = gm1_d * (qinf_d[3] - 0.5 * ri * (qinf_d[1] * qinf_d[1] + qinf_d[2] * qinf_d[2]));
    vol2 = ri * (qinf_d[1] * dy - qinf_d[2] * dx);
    mu = (*adt1) * eps_d;
    f = 0.5 * (vol1 * q1[0] + vol2 * qinf_d[0]) + mu * (q1[0] - qinf_d[0]);
    res1[0] += f;
    f = 0.5 * (vol1 * q1[1] + p1 * dy + vol2 * qinf_d[1] + p2 * dy) +
        mu * (q1[1] - qinf_d[1]);
    res1[1] += f;
    f = 0.5 * (vol1 * q1[2] - p1 * dx + vol2 * qinf_d[2] - p2 * dx) +
        mu * (q1[2] - qinf_d[2]);
    res1[2] += f;
    f = 0.5 * (vol1 * (q1[3] + p1) + vol2 * (qinf_d[3] + p2)) +
        mu * (q1[3] - qinf_d[3]);
    res1[3] += f;
  }
}

__device__ __forceinline__ void update(const double *qold, double *q,
                                       double *res, const double *adt,
                                       double *rms, double *maxerr,
                                       const int *idx, int *errloc) {
  double del, adti;
  adti = 1.0 / (*adt);
  for (int n = 0; n < 4; n++) {
    del = adti * res[n];
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

}  // namespace op2_k

// FP32 user kernels of the mixed-precision mode: the reference's
// single-precision airfoil kernels (apps/c/airfoil/airfoil_plain/sp/
// adt_calc.h, res_calc.h, bres_calc.h) on FP32 copies of q and x.
namespace op2_kf {

__device__ __forceinline__ void adt_calc(const float *x1, const float *x2,
                                         const float *x3, const float *x4,
                                         const float *q, float *adt) {
  float dx, dy, ri, u, v, c;
  ri = 1.0f / q[0];
  u = ri * q[1];
  v = ri * q[2];
  c = sqrtf(gam_f * gm1_f * (ri * q[3] - 0.5f * (u * u + v * v)));
  dx = x2[0] - x1[0];
  dy = x2[1] - x1[1];
  *adt = fabsf(u * dy - v * dx) + c * sqrtf(dx * dx + dy * dy);
  dx = x3[0] - x2[0];
  dy = x3[1] - x2[1];
  *adt += fabsf(u * dy - v * dx) + c * sqrtf(dx * dx + dy * dy);
  dx = x4[0] - x3[0];
  dy = x4[1] - x3[1];
  *adt += fabsf(u * dy - v * dx) + c * sqrtf(dx * dx + dy * dy);
  dx = x1[0] - x4[0];
  dy = x1[1] - x4[1];
  *adt += fabsf(u * dy - v * dx) + c * sqrtf(dx * dx + dy * dy);
  *adt = (*adt) / cfl_f;
}

__device__ __forceinline__ void res_calc(const float *x1, const float *x2,
                                         const float *q1, const float *q2,
                                         const float *adt1, const float *adt2,
                                         float *res1, float *res2) {
  float dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0f / q1[0];
  p1 = gm1_f * (q1[3] - 0.5f * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0f / q2[0];
  p2 = gm1_f * (q2[3] - 0.5f * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  vol2 = ri * (q2[1] * dy - q2[2] * dx);
  mu = 0.5f * ((*adt1) + (*adt2)) * eps_f;
  f = 0.5f * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  res1[0] += f;
  res2[0] -= f;
  f = 0.5f * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) + mu * (q1[1] - q2[1]);
  res1[1] += f;
  res2[1] -= f;
  f = 0.5f * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) + mu * (q1[2] - q2[2]);
  res1[2] += f;
  res2[2] -= f;
  f = 0.5f * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
  res1[3] += f;
  res2[3] -= f;
}

__device__ __forceinline__ void bres_calc(const float *x1, const float *x2,
                                          const float *q1, const float *adt1,
                                          float *res1, const int *bound) {
  float dx, dy, mu, ri, p1, vol1, p2, vol2, f;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0f / q1[0];
  p1 = gm1_f * (q1[3] - 0.5f * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  if (*bound == 1) {
    res1[1] += +p1 * dy;
    res1[2] += -p1 * dx;
  } else {
    vol1 = ri * (q1[1] * dy - q1[2] * dx);
    ri = 1.0f / qinf_f[0];
    p2 = gm1_f * (qinf_f[3] - 0.5f * ri * (qinf_f[1] * qinf_f[1] + qinf_f[2] * qinf_f[2]));
    vol2 = ri * (qinf_f[1] * dy - qinf_f[2] * dx);
    mu = (*adt1) * eps_f;
    f = 0.5f * (vol1 * q1[0] + vol2 * qinf_f[0]) + mu * (q1[0] - qinf_f[0]);
    res1[0] += f;
    f = 0.5f * (vol1 * q1[1] + p1 * dy + vol2 * qinf_f[1] + p2 * dy) + mu * (q1[1] - qinf_f[1]);
    res1[1] += f;
    f = 0.5f * (vol1 * q1[2] - p1 * dx + vol2 * qinf_f[2] - p2 * dx) + mu * (q1[2] - qinf_f[2]);
    res1[2] += f;
    f = 0.5f * (vol1 * (q1[3] + p1) + vol2 * (qinf_f[3] + p2)) + mu * (q1[3] - qinf_f[3]);
    res1[3] += f;
  }
}

// The increment f of res_calc (SP), like op2_k::res_flux.
__device__ __forceinline__ void res_flux(const float *x1, const float *x2,
                                         const float *q1, const float *q2,
                                         float adt1, float adt2, float *f) {
  float dx, dy, mu, ri, p1, vol1, p2, vol2;
  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];
  ri = 1.0f / q1[0];
  p1 = gm1_f * (q1[3] - 0.5f * ri * (q1[1] * q1[1] + q1[2] * q1[2]));
  vol1 = ri * (q1[1] * dy - q1[2] * dx);
  ri = 1.0f / q2[0];
  p2 = gm1_f * (q2[3] - 0.5f * ri * (q2[1] * q2[1] + q2[2] * q2[2]));
  vol2 = ri * (q2[1] * dy - q2[2] * dx);
  mu = 0.5f * (adt1 + adt2) * eps_f;
  f[0] = 0.5f * (vol1 * q1[0] + vol2 * q2[0]) + mu * (q1[0] - q2[0]);
  f[1] = 0.5f * (vol1 * q1[1] + p1 * dy + vol2 * q2[1] + p2 * dy) + mu * (q1[1] - q2[1]);
  f[2] = 0.5f * (vol1 * q1[2] - p1 * dx + vol2 * q2[2] - p2 * dx) + mu * (q1[2] - q2[2]);
  f[3] = 0.5f * (vol1 * (q1[3] + p1) + vol2 * (q2[3] + p2)) + mu * (q1[3] - q2[3]);
}

}  // namespace op2_kf

// ---------------------------------------------------------------------------
// dat access: AoS (element-contiguous, 16-byte vector loads) or SoA
// (component rows `stride` elements apart)
// ---------------------------------------------------------------------------
namespace op2_dev {

template <bool SOA>
__device__ __forceinline__ void ld2(const double *__restrict p, int n, int s,
                                    double *v) {
  if (SOA) {
    v[0] = p[n];
    v[1] = p[(size_t)s + n];
  } else {
    double2 a = *reinterpret_cast<const double2 *>(p + 2 * (size_t)n);
    v[0] = a.x;
    v[1] = a.y;
  }
}

template <bool SOA>
__device__ __forceinline__ void ld4(const double *__restrict p, int n, int s,
                                    double *v) {
  if (SOA) {
    v[0] = p[n];
    v[1] = p[(size_t)s + n];
    v[2] = p[2 * (size_t)s + n];
    v[3] = p[3 * (size_t)s + n];
  } else {
    double2 a = *reinterpret_cast<const double2 *>(p + 4 * (size_t)n);
    double2 b = *reinterpret_cast<const double2 *>(p + 4 * (size_t)n + 2);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  }
}

template <bool SOA>
__device__ __forceinline__ void st4(double *__restrict p, int n, int s,
                                    const double *v) {
  if (SOA) {
    p[n] = v[0];
    p[(size_t)s + n] = v[1];
    p[2 * (size_t)s + n] = v[2];
    p[3 * (size_t)s + n] = v[3];
  } else {
    *reinterpret_cast<double2 *>(p + 4 * (size_t)n) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(p + 4 * (size_t)n + 2) = make_double2(v[2], v[3]);
  }
}

// Non-temporal forms of the element-indexed (single-use) accesses, for
// loops whose streams exceed the last-level cache
typedef double nt_double2 __attribute__((ext_vector_type(2)));
typedef int nt_int4 __attribute__((ext_vector_type(4)));

template <bool SOA>
__device__ __forceinline__ void ld4_nt(const double *__restrict p, int n, int s,
                                       double *v) {
  if (SOA) {
    v[0] = __builtin_nontemporal_load(p + n);
    v[1] = __builtin_nontemporal_load(p + (size_t)s + n);
    v[2] = __builtin_nontemporal_load(p + 2 * (size_t)s + n);
    v[3] = __builtin_nontemporal_load(p + 3 * (size_t)s + n);
  } else {
    const nt_double2 a =
        __builtin_nontemporal_load(reinterpret_cast<const nt_double2 *>(p) + 2 * (size_t)n);
    const nt_double2 b =
        __builtin_nontemporal_load(reinterpret_cast<const nt_double2 *>(p) + 2 * (size_t)n + 1);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  }
}

__device__ __forceinline__ int4 ld_int4_nt(const int *__restrict p, int n) {
  const nt_int4 t = __builtin_nontemporal_load(reinterpret_cast<const nt_int4 *>(p) + n);
  return make_int4(t.x, t.y, t.z, t.w);
}

template <bool SOA>
__device__ __forceinline__ void add4(double *__restrict p, int n, int s,
                                     const double *v) {
  double o[4];
  ld4<SOA>(p, n, s, o);
  o[0] += v[0];
  o[1] += v[1];
  o[2] += v[2];
  o[3] += v[3];
  st4<SOA>(p, n, s, o);
}

template <bool SOA>
__device__ __forceinline__ void atomic_add4(double *p, int n, int s,
                                            const double *v) {
  size_t st = SOA ? (size_t)s : 1, base = SOA ? (size_t)n : 4 * (size_t)n;
  unsafeAtomicAdd(p + base, v[0]);
  unsafeAtomicAdd(p + base + st, v[1]);
  unsafeAtomicAdd(p + base + 2 * st, v[2]);
  unsafeAtomicAdd(p + base + 3 * st, v[3]);
}

inline int grid_for(long long n, int bs) { return n > 0 ? (int)((n + bs - 1) / bs) : 0; }

// Workgroups are dispatched round-robin over the XCDs (blockIdx % nxcd).
// Logical block ids are assigned so that each XCD runs one contiguous range
// of them: neighbours in a locally numbered mesh then stay in one XCD's L2.
// The map is a bijection of [0, gridDim.x).
__device__ __forceinline__ int xcd_block(int nxcd) {
  const int b = blockIdx.x, n = gridDim.x;
  if (nxcd <= 1 || n < nxcd) return b;
  const int q = n / nxcd, r = n % nxcd;
  const int x = b % nxcd, i = b / nxcd;
  const int start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return start + i;
}

constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2

// The value held by lane J of this lane's quad (lanes 4q..4q+3 of the wave),
// by a DPP quad_perm(J, J, J, J) move: a VALU operation, no LDS traffic.
// Every lane of the quad must be active.
template <int J>
__device__ __forceinline__ double quad_bcast(double v) {
  constexpr int ctrl = J * 0x55;
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int J>
__device__ __forceinline__ float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), J * 0x55, 0xF, 0xF, false));
}
template <int J>
__device__ __forceinline__ int quad_bcast(int v) {
  return __builtin_amdgcn_mov_dpp(v, J * 0x55, 0xF, 0xF, false);
}

// ---------------------------------------------------------------------------
// Owner-compute res_calc (the owner partition plan): per block its instance
// records, owned cells and per-cell increment lists
// ---------------------------------------------------------------------------
struct OwnerArgs {
  const int4 *inst;  // per instance: edge nodes (x, y), edge cells (z, w)
  // compact form: per block {cell base, node base} (node base -1: the block's
  // spans do not fit 16 bits and it reads `inst`), per instance the offsets
  // {n0 | n1 << 16, c0 | c1 << 16} from those bases
  const int2 *base;
  const uint2 *inst8;
  const int *elem_offs, *owned_offs, *owned, *csr_offs;
  const unsigned short *csr;
  const ushort4 *csr4;  // per owned cell: up to four increments, 0xFFFF = none
  int max_e;
};

// Instance j of a block with bases bb: from the 8-byte record when the
// block has one (node base >= 0, uniform over the workgroup), else the
// 16-byte record.
__device__ __forceinline__ int4 owner_record(const OwnerArgs &P, int2 bb, int j) {
  if (bb.y >= 0) {
    const uint2 r = P.inst8[j];
    return make_int4(bb.y + (int)(r.x & 0xFFFFu), bb.y + (int)(r.x >> 16),
                     bb.x + (int)(r.y & 0xFFFFu), bb.x + (int)(r.y >> 16));
  }
  return P.inst[j];
}

// Adds one owned cell's increments to acc in list order: +f for ecell[0]
// (even entries), -f for ecell[1] (odd).
template <class T>
__device__ __forceinline__ void owner_apply(T *acc, const T *lf, int ent) {
  const T *f = lf + 4 * (ent >> 1);
  if (ent & 1) {
    for (int d = 0; d < 4; ++d) acc[d] -= f[d];
  } else {
    for (int d = 0; d < 4; ++d) acc[d] += f[d];
  }
}

// Owner-compute res_calc of block b:
//  1. one flux per edge touching the block (edges on the block boundary are
//     evaluated by both neighbouring blocks), from global memory, into LDS;
//  2. each owned cell adds its edges' fluxes to res in ascending edge order
//     (+f where it is ecell[0], -f where it is ecell[1]), as the sequential
//     executor does, so res is bit-identical to it.  zero_acc: res is known
//     to hold +0.0 in every owned cell (the update before it zeroed it), so
//     the sums start from +0.0 instead of reading it: the same bits.
// lf: LDS for the block's fluxes, [max_e][4].
template <bool SOA, bool C4>
__device__ __forceinline__ void owner_res_block(int b, const double *__restrict x,
                                                const double *__restrict q,
                                                const double *__restrict adt,
                                                double *__restrict res, const OwnerArgs &P,
                                                int sx, int sq, double *lf,
                                                bool zero_acc = false) {
  const int e0 = P.elem_offs[b], ne = P.elem_offs[b + 1] - e0;
  const int2 bb = P.base[b];
  for (int j = threadIdx.x; j < ne; j += blockDim.x) {
    const int4 r = owner_record(P, bb, e0 + j);
    double x1[2], x2[2], q1[4], q2[4];
    ld2<SOA>(x, r.x, sx, x1);
    ld2<SOA>(x, r.y, sx, x2);
    ld4<SOA>(q, r.z, sq, q1);
    ld4<SOA>(q, r.w, sq, q2);
    op2_k::res_flux(x1, x2, q1, q2, adt[r.z], adt[r.w], lf + 4 * j);
  }
  __syncthreads();
  const int o0 = P.owned_offs[b], no = P.owned_offs[b + 1] - o0;
  for (int i = threadIdx.x; i < no; i += blockDim.x) {
    const int cell = P.owned[o0 + i];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (!zero_acc) ld4<SOA>(res, cell, sq, acc);
    if (C4) {
      const ushort4 e = P.csr4[o0 + i];
      if (e.x != 0xFFFF) owner_apply(acc, lf, e.x);
      if (e.y != 0xFFFF) owner_apply(acc, lf, e.y);
      if (e.z != 0xFFFF) owner_apply(acc, lf, e.z);
      if (e.w != 0xFFFF) owner_apply(acc, lf, e.w);
    } else {
      for (int k = P.csr_offs[o0 + i]; k < P.csr_offs[o0 + i + 1]; ++k)
        owner_apply(acc, lf, (int)P.csr[k]);
    }
    st4<SOA>(res, cell, sq, acc);
  }
}

// res_calc and bres_calc fused into one owner-compute launch (loop-chain
// fusion inside a captured step): each owned cell's fixed-width record lists
// its interior increments, then its boundary edges as 0x8000 | j, j indexing
// the block's list of boundary edges.  The interior fluxes are summed first,
// in edge order, then the boundary fluxes, added by bres_calc itself into the
// running sum in boundary-edge order: the order of res_calc then bres_calc in
// the sequential executor, so res is bit-identical to it.
struct BoundaryArgs {
  const ushort4 *rec;   // per owned cell: up to four entries, 0xFFFF = none
  const int *offs;      // [nblocks+1] per block: its boundary edges' start
  const int *edges;     // boundary edge of each (block, j)
  const int2 *pbedge;   // boundary edge -> its two nodes
  const int *bound;     // boundary edge -> 1 wall, else far field
  // ensembles (QI): boundary edge -> its member, and the members' free
  // streams [member][4]
  const int *bmesh;
  const double *qinf;
};

// QI: each boundary edge with its member's free stream (B.bmesh, B.qinf),
// else the program's qinf.
template <bool SOA, bool QI = false>
__device__ __forceinline__ void owner_res_bres_block(int b, const double *__restrict x,
                                                     const double *__restrict q,
                                                     const double *__restrict adt,
                                                     double *__restrict res,
                                                     const OwnerArgs &P, const BoundaryArgs &B,
                                                     int sx, int sq, double *lf,
                                                     bool zero_acc = false) {
  const int e0 = P.elem_offs[b], ne = P.elem_offs[b + 1] - e0;
  const int2 bb = P.base[b];
  for (int j = threadIdx.x; j < ne; j += blockDim.x) {
    const int4 r = owner_record(P, bb, e0 + j);
    double x1[2], x2[2], q1[4], q2[4];
    ld2<SOA>(x, r.x, sx, x1);
    ld2<SOA>(x, r.y, sx, x2);
    ld4<SOA>(q, r.z, sq, q1);
    ld4<SOA>(q, r.w, sq, q2);
    op2_k::res_flux(x1, x2, q1, q2, adt[r.z], adt[r.w], lf + 4 * j);
  }
  __syncthreads();
  const int o0 = P.owned_offs[b], no = P.owned_offs[b + 1] - o0;
  const int bb0 = B.offs[b];
  for (int i = threadIdx.x; i < no; i += blockDim.x) {
    const int cell = P.owned[o0 + i];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (!zero_acc) ld4<SOA>(res, cell, sq, acc);
    const ushort4 e = B.rec[o0 + i];
    const unsigned short ent[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (ent[k] == 0xFFFF) continue;
      if (ent[k] & 0x8000) {
        const int be = B.edges[bb0 + (ent[k] & 0x7FFF)];
        const int2 bn = B.pbedge[be];
        double x1[2], x2[2], q1[4];
        ld2<SOA>(x, bn.x, sx, x1);
        ld2<SOA>(x, bn.y, sx, x2);
        ld4<SOA>(q, cell, sq, q1);
        const double a1 = adt[cell];
        const int bnd = B.bound[be];
        if constexpr (QI)
          op2_k::bres_calc_q(x1, x2, q1, &a1, acc, &bnd, B.qinf + 4 * (size_t)B.bmesh[be]);
        else
          op2_k::bres_calc(x1, x2, q1, &a1, acc, &bnd);
      } else {
        owner_apply(acc, lf, (int)ent[k]);
      }
    }
    st4<SOA>(res, cell, sq, acc);
  }
}

}  // namespace op2_dev
