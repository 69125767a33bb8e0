// Batched airfoil ensembles (include/op2amd_ensemble.h).
//
// Members are concatenated into one mesh; every loop of the time step is one
// launch over all members, except in the looped baseline, which launches each
// loop once per member over that member's ranges.  res_calc and bres_calc use
// the owner-compute plan (each cell sums its edges' fluxes in edge order), so
// results do not depend on which launch structure computed them.
#include "op2amd_ensemble.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "airfoil_device.h"
#include "op2h_plan_build.h"

#define ECHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip ensemble: %s failed at %s:%d: %s\n", #x, __FILE__,      \
             __LINE__, hipGetErrorString(e_));                               \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace {

using namespace op2_dev;

constexpr int kTile = 1024;  // cells per update tile (one per thread)
constexpr int kPatch = 128;  // cells per owner patch of res_calc (one per thread)

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

double now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---- kernels ------------------------------------------------------------

__global__ void k_ens_copy(const double2 *__restrict q, double2 *__restrict qold,
                           long long lo2, long long hi2) {
  long long i = lo2 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < hi2) qold[i] = q[i];
}

__global__ void k_ens_adt(const double *__restrict x, const double *__restrict q,
                          double *__restrict adt, const int4 *__restrict pcell,
                          int lo, int hi) {
  int c = lo + xcd_block(kXcds) * blockDim.x + threadIdx.x;
  if (c >= hi) return;
  int4 cn = pcell[c];
  double x1[2], x2[2], x3[2], x4[2], qc[4], a;
  ld2<false>(x, cn.x, 0, x1);
  ld2<false>(x, cn.y, 0, x2);
  ld2<false>(x, cn.z, 0, x3);
  ld2<false>(x, cn.w, 0, x4);
  ld4<false>(q, c, 0, qc);
  op2_k::adt_calc(x1, x2, x3, x4, qc, &a);
  adt[c] = a;
}

// res_calc over blocks [blo, blo + gridDim.x) of the members' owner
// partitions: each edge evaluated once per patch it touches, each cell's
// increments summed in edge order (op2_dev::owner_res_block), so every member
// gets its sequential executor's bits.  LDS: fluxes [max_e][4].
__global__ void __launch_bounds__(1024)
    k_ens_res_owner(const double *__restrict x, const double *__restrict q,
                    const double *__restrict adt, double *__restrict res, OwnerArgs P,
                    int blo) {
  extern __shared__ double lf[];
  owner_res_block<false, true>(blo + xcd_block(kXcds), x, q, adt, res, P, 0, 0, lf);
}

// bres_calc (bres_calc.h:1-35) with the member's own free stream `qi`
__device__ __forceinline__ void bres_calc_qi(const double *x1, const double *x2,
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

// one thread per boundary cell; its boundary edges in edge order
__global__ void k_ens_bres(const double *__restrict x, const double *__restrict q,
                           const double *__restrict adt, double *res,
                           const int2 *__restrict pbedge, const int *__restrict bound,
                           const int *__restrict bmesh, const double *__restrict qinf,
                           const int *__restrict offsets, const int *__restrict entries,
                           const int *__restrict targets, int tlo, int thi) {
  int t = tlo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= thi) return;
  int cell = targets[t];
  double acc[4], q1[4];
  ld4<false>(res, cell, 0, acc);
  ld4<false>(q, cell, 0, q1);
  double a1 = adt[cell];
  for (int k = offsets[t]; k < offsets[t + 1]; ++k) {
    int e = entries[k];
    int2 bn = pbedge[e];
    double x1[2], x2[2], r[4] = {0.0, 0.0, 0.0, 0.0};
    ld2<false>(x, bn.x, 0, x1);
    ld2<false>(x, bn.y, 0, x2);
    int bd = bound[e];
    bres_calc_qi(x1, x2, q1, &a1, r, &bd, qinf + 4 * (size_t)bmesh[e]);
    for (int d = 0; d < 4; ++d) acc[d] += r[d];
  }
  st4<false>(res, cell, 0, acc);
}

struct MaxLoc {
  double v;
  int loc;
};

__device__ __forceinline__ MaxLoc maxloc_combine(MaxLoc a, MaxLoc b) {
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

// Persistent update over the tile table: workgroup w processes tiles
// w, w + gridDim.x, ... of [tlo, thi); a tile is <= kTile cells of one member.
__global__ void __launch_bounds__(kTile)
    k_ens_update(const double *__restrict qold, double *__restrict q,
                 double *__restrict res, const double *__restrict adt,
                 const int *__restrict tile_start, const int *__restrict tile_len,
                 const int *__restrict tile_base, int tlo, int thi,
                 double *__restrict part_sum, double *__restrict part_max,
                 int *__restrict part_loc) {
  __shared__ double s_sum[kTile / 64];
  __shared__ double s_max[kTile / 64];
  __shared__ int s_loc[kTile / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = tlo + blockIdx.x; t < thi; t += gridDim.x) {
    double rms = 0.0;
    MaxLoc ml{-INFINITY, 0x7fffffff};
    if ((int)threadIdx.x < tile_len[t]) {
      int c = tile_start[t] + threadIdx.x;
      int local = c - tile_base[t];  // index within the member
      double qo[4], qn[4], r[4], a = adt[c], mx = ml.v;
      int loc = ml.loc;
      ld4<false>(qold, c, 0, qo);
      ld4<false>(res, c, 0, r);
      op2_k::update(qo, qn, r, &a, &rms, &mx, &local, &loc);
      ml.v = mx;
      ml.loc = loc;
      st4<false>(q, c, 0, qn);
      st4<false>(res, c, 0, r);
    }
    rms = wave_sum(rms);
    ml = wave_maxloc(ml);
    if (lane == 0) {
      s_sum[wave] = rms;
      s_max[wave] = ml.v;
      s_loc[wave] = ml.loc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double bs = 0.0;
      MaxLoc bm{-INFINITY, 0x7fffffff};
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        bs += s_sum[w];
        bm = maxloc_combine(bm, MaxLoc{s_max[w], s_loc[w]});
      }
      part_sum[t] = bs;
      part_max[t] = bm.v;
      part_loc[t] = bm.loc;
    }
    __syncthreads();
  }
}

// One workgroup per member: its tiles' partials in tile order.
__global__ void __launch_bounds__(256)
    k_ens_finish(const double *__restrict part_sum, const double *__restrict part_max,
                 const int *__restrict part_loc, const int *__restrict mesh_tile_off,
                 int mlo, double *__restrict slots) {
  __shared__ double s_sum[4];
  __shared__ double s_max[4];
  __shared__ int s_loc[4];
  const int m = mlo + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double tot = 0.0;
  MaxLoc tm{-INFINITY, 0x7fffffff};
  for (int t = mesh_tile_off[m] + tid; t < mesh_tile_off[m + 1]; t += 256) {
    tot += part_sum[t];
    tm = maxloc_combine(tm, MaxLoc{part_max[t], part_loc[t]});
  }
  tot = wave_sum(tot);
  tm = wave_maxloc(tm);
  if (lane == 0) {
    s_sum[wave] = tot;
    s_max[wave] = tm.v;
    s_loc[wave] = tm.loc;
  }
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    MaxLoc mm{-INFINITY, 0x7fffffff};
    for (int w = 0; w < 4; ++w) {
      t += s_sum[w];
      mm = maxloc_combine(mm, MaxLoc{s_max[w], s_loc[w]});
    }
    slots[3 * (size_t)m] = t;
    slots[3 * (size_t)m + 1] = mm.v;
    slots[3 * (size_t)m + 2] = (double)mm.loc;
  }
}

// history[record][k] = slots[k]; one workgroup; the counter advances once.
__global__ void k_ens_history(const double *__restrict slots, int nwords,
                              double *__restrict hist, unsigned long long *count,
                              int capacity) {
  unsigned long long c = *count;
  size_t base = (size_t)(c % (unsigned long long)capacity) * nwords;
  for (int k = threadIdx.x; k < nwords; k += blockDim.x) hist[base + k] = slots[k];
  __syncthreads();
  if (threadIdx.x == 0) *count = c + 1;
}

// ---- host -----------------------------------------------------------------

template <class T> T *dev_upload(const std::vector<T> &v, std::vector<void *> &keep) {
  void *p = nullptr;
  ECHECK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty()) ECHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  keep.push_back(p);
  return (T *)p;
}

template <class T> T *dev_alloc(size_t n, std::vector<void *> &keep) {
  void *p = nullptr;
  ECHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  ECHECK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  ECHECK(hipDeviceSynchronize());  // done before the non-blocking stream uses it
  keep.push_back(p);
  return (T *)p;
}

}  // namespace

extern "C" int op2amd_ensemble_describe(const op2amd_ensemble_spec *spec,
                                        op2amd_ensemble_member *members) {
  if (!spec || !members || spec->nmesh < 1 || spec->min_I < 2 || spec->min_J < 2 ||
      spec->max_I < spec->min_I)
    return 1;
  SplitMix64 rng(spec->seed);
  const double smax = (double)spec->max_I / spec->min_I;
  // constants of airfoil.cpp:149-159
  const double gam = 1.4f, gm1 = gam - 1.0f, p = 1.0f, r = 1.0f;
  for (int m = 0; m < spec->nmesh; ++m) {
    double s = 1.0 + (smax - 1.0) * rng.uniform();
    double um = rng.uniform(), ua = rng.uniform();
    op2amd_ensemble_member &mb = members[m];
    mb.I = std::max(2, (int)std::lround(spec->min_I * s));
    mb.J = std::max(2, (int)std::lround(spec->min_J * s));
    mb.ncell = mb.I * mb.J;
    mb.mach = spec->mach_lo + (spec->mach_hi - spec->mach_lo) * um;
    mb.alpha_deg = spec->alpha_lo_deg + (spec->alpha_hi_deg - spec->alpha_lo_deg) * ua;
    double a = mb.alpha_deg * (M_PI / 180.0);
    double u = sqrt(gam * p / r) * mb.mach;
    double e = p / (r * gm1) + 0.5f * u * u;
    mb.qinf[0] = r;
    mb.qinf[1] = r * u * cos(a);
    mb.qinf[2] = r * u * sin(a);
    mb.qinf[3] = r * e;
  }
  return 0;
}

extern "C" int op2amd_ensemble_run(const op2amd_ensemble_spec *spec,
                                   const op2amd_ensemble_opts *o,
                                   op2amd_ensemble_result *res) {
  if (!spec || !o || !res || o->niter < 1 || o->warmup < 0 || o->warmup >= o->niter)
    return 1;
  double t_setup = now();
  const int M = spec->nmesh;
  std::vector<op2amd_ensemble_member> mem(M);
  if (op2amd_ensemble_describe(spec, mem.data())) return 1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    printf("op2hip ensemble: no HIP device available\n");
    return 2;
  }
  ECHECK(hipSetDevice(0));

  // ---- concatenate the members ----
  std::vector<int> noff(M + 1, 0), coff(M + 1, 0), eoff(M + 1, 0), boff(M + 1, 0);
  std::vector<double> x;
  std::vector<int4> cell;
  std::vector<int2> edge, ecell, bedge;
  std::vector<int> becell, bound, bmesh;
  for (int m = 0; m < M; ++m) {
    op2amd_mesh mesh;
    if (op2amd_mesh_generate(mem[m].I, mem[m].J, spec->seed + 1 + (uint64_t)m,
                             spec->ordering, &mesh))
      return 3;
    const int n0 = noff[m], c0 = coff[m];
    x.insert(x.end(), mesh.x, mesh.x + 2 * (size_t)mesh.nnode);
    for (int c = 0; c < mesh.ncell; ++c)
      cell.push_back(make_int4(mesh.cell[4 * c] + n0, mesh.cell[4 * c + 1] + n0,
                               mesh.cell[4 * c + 2] + n0, mesh.cell[4 * c + 3] + n0));
    for (int e = 0; e < mesh.nedge; ++e) {
      edge.push_back(make_int2(mesh.edge[2 * e] + n0, mesh.edge[2 * e + 1] + n0));
      ecell.push_back(make_int2(mesh.ecell[2 * e] + c0, mesh.ecell[2 * e + 1] + c0));
    }
    for (int b = 0; b < mesh.nbedge; ++b) {
      bedge.push_back(make_int2(mesh.bedge[2 * b] + n0, mesh.bedge[2 * b + 1] + n0));
      becell.push_back(mesh.becell[b] + c0);
      bound.push_back(mesh.bound[b]);
      bmesh.push_back(m);
    }
    noff[m + 1] = n0 + mesh.nnode;
    coff[m + 1] = c0 + mesh.ncell;
    eoff[m + 1] = eoff[m] + mesh.nedge;
    boff[m + 1] = boff[m] + mesh.nbedge;
    op2amd_mesh_free(&mesh);
  }
  const int Nc = coff[M], Ne = eoff[M], Nb = boff[M];
  res->total_cells = Nc;
  res->total_edges = Ne;

  // res_calc: each member's owner partition (patches of kPatch cells that
  // never straddle two members), concatenated; 8-byte instance records where a
  // patch's cells and nodes span at most 65535 indices
  std::vector<int> mblk(M + 1, 0), owned_offs{0}, owned, elem_offs{0};
  std::vector<int4> inst;
  std::vector<uint2> inst8;
  std::vector<int2> pbase;
  std::vector<ushort4> csr4;
  int max_e = 0;
  {
    std::vector<int> lmap;
    const int inc[2] = {0, 1};
    for (int m = 0; m < M; ++m) {
      const int c0 = coff[m], e0 = eoff[m], nel = eoff[m + 1] - e0;
      lmap.resize(2 * (size_t)nel);
      for (int e = 0; e < nel; ++e) {
        lmap[2 * (size_t)e] = ecell[e0 + e].x - c0;
        lmap[2 * (size_t)e + 1] = ecell[e0 + e].y - c0;
      }
      op2h::OwnerLists L;
      std::string err;
      if (!op2h::owner_lists(coff[m + 1] - c0, nel, 2, lmap.data(), 2, inc, kPatch, L, err)) {
        printf("op2hip ensemble: member %d: %s\n", m, err.c_str());
        return 4;
      }
      if (L.max_incs > 4) {
        printf("op2hip ensemble: a cell has more than 4 edges\n");
        return 4;
      }
      max_e = std::max(max_e, L.max_elems);
      for (int b = 0; b < L.nblocks; ++b) {
        const size_t first = inst.size();
        int cmin = INT_MAX, cmax = -1, nmin = INT_MAX, nmax = -1;
        for (int j = L.elem_offs[b]; j < L.elem_offs[b + 1]; ++j) {
          const int e = e0 + L.elems[j];
          const int4 r = make_int4(edge[e].x, edge[e].y, ecell[e].x, ecell[e].y);
          inst.push_back(r);
          nmin = std::min(nmin, std::min(r.x, r.y));
          nmax = std::max(nmax, std::max(r.x, r.y));
          cmin = std::min(cmin, std::min(r.z, r.w));
          cmax = std::max(cmax, std::max(r.z, r.w));
        }
        const bool fits = inst.size() > first && cmax - cmin <= 0xFFFF && nmax - nmin <= 0xFFFF;
        pbase.push_back(fits ? make_int2(cmin, nmin) : make_int2(0, -1));
        for (size_t j = first; j < inst.size(); ++j) {
          const int4 r = inst[j];
          inst8.push_back(fits ? make_uint2((unsigned)(r.x - nmin) | ((unsigned)(r.y - nmin) << 16),
                                            (unsigned)(r.z - cmin) | ((unsigned)(r.w - cmin) << 16))
                               : make_uint2(0u, 0u));
        }
        elem_offs.push_back((int)inst.size());
        for (int i = L.owned_offs[b]; i < L.owned_offs[b + 1]; ++i) {
          owned.push_back(c0 + L.owned[i]);
          unsigned short v[4] = {0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF};
          for (int k = L.csr_offs[i]; k < L.csr_offs[i + 1]; ++k) v[k - L.csr_offs[i]] = L.csr[k];
          csr4.push_back(make_ushort4(v[0], v[1], v[2], v[3]));
        }
        owned_offs.push_back((int)owned.size());
      }
      mblk[m + 1] = mblk[m] + L.nblocks;
    }
  }
  // bres: boundary cells in cell order with their bedges in bedge order
  std::vector<int> bcount(Nc, 0);
  for (int b = 0; b < Nb; ++b) bcount[becell[b]]++;
  std::vector<int> btargets, boffsets{0}, bentries(Nb);
  std::vector<int> bpos(Nc, -1);
  for (int c = 0; c < Nc; ++c)
    if (bcount[c]) {
      bpos[c] = (int)btargets.size();
      btargets.push_back(c);
      boffsets.push_back(boffsets.back() + bcount[c]);
    }
  {
    std::vector<int> fill(boffsets.begin(), boffsets.end() - 1);
    for (int b = 0; b < Nb; ++b) bentries[fill[bpos[becell[b]]]++] = b;
  }
  std::vector<int> mesh_btarget_off(M + 1, 0);
  for (int m = 0; m <= M; ++m)
    mesh_btarget_off[m] = (int)(std::lower_bound(btargets.begin(), btargets.end(), coff[m]) -
                                btargets.begin());
  // update tiles: <= kTile cells, never straddling two members
  std::vector<int> tile_start, tile_len, tile_base, mesh_tile_off{0};
  for (int m = 0; m < M; ++m) {
    for (int c = coff[m]; c < coff[m + 1]; c += kTile) {
      tile_start.push_back(c);
      tile_len.push_back(std::min(kTile, coff[m + 1] - c));
      tile_base.push_back(coff[m]);
    }
    mesh_tile_off.push_back((int)tile_start.size());
  }
  const int ntiles = (int)tile_start.size();

  // ---- device data ----
  std::vector<void *> keep;
  std::vector<double> q0(4 * (size_t)Nc), qinf(4 * (size_t)M);
  for (int m = 0; m < M; ++m)
    for (int k = 0; k < 4; ++k) qinf[4 * m + k] = mem[m].qinf[k];
  for (int m = 0; m < M; ++m)
    for (int c = coff[m]; c < coff[m + 1]; ++c)
      for (int k = 0; k < 4; ++k) q0[4 * (size_t)c + k] = mem[m].qinf[k];
  double *d_x = dev_upload(x, keep);
  double *d_q = dev_upload(q0, keep);
  double *d_qold = dev_upload(q0, keep);
  double *d_res = dev_alloc<double>(4 * (size_t)Nc, keep);
  double *d_adt = dev_alloc<double>(Nc, keep);
  int4 *d_cell = dev_upload(cell, keep);
  OwnerArgs OA{};
  OA.inst = dev_upload(inst, keep);
  OA.base = dev_upload(pbase, keep);
  OA.inst8 = dev_upload(inst8, keep);
  OA.elem_offs = dev_upload(elem_offs, keep);
  OA.owned_offs = dev_upload(owned_offs, keep);
  OA.owned = dev_upload(owned, keep);
  OA.csr4 = dev_upload(csr4, keep);
  OA.max_e = max_e;
  const size_t res_lds = sizeof(double) * 4 * (size_t)max_e;
  int2 *d_bedge = dev_upload(bedge, keep);
  int *d_bound = dev_upload(bound, keep);
  int *d_bmesh = dev_upload(bmesh, keep);
  double *d_qinf = dev_upload(qinf, keep);
  int *d_btargets = dev_upload(btargets, keep);
  int *d_boffsets = dev_upload(boffsets, keep);
  int *d_bentries = dev_upload(bentries, keep);
  int *d_tstart = dev_upload(tile_start, keep);
  int *d_tlen = dev_upload(tile_len, keep);
  int *d_tbase = dev_upload(tile_base, keep);
  int *d_mtoff = dev_upload(mesh_tile_off, keep);
  double *d_psum = dev_alloc<double>(ntiles, keep);
  double *d_pmax = dev_alloc<double>(ntiles, keep);
  int *d_ploc = dev_alloc<int>(ntiles, keep);
  double *d_slots = dev_alloc<double>(3 * (size_t)M, keep);
  const int hist_words = 3 * M;
  const int capacity = o->niter;
  double *d_hist = dev_alloc<double>((size_t)hist_words * capacity, keep);
  unsigned long long *d_count = dev_alloc<unsigned long long>(1, keep);
  {
    // constants of airfoil.cpp:149-152 in this translation unit's copies
    double gam = 1.4f, gm1 = gam - 1.0f, cfl = 0.9f, eps = 0.05f;
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(gam_d), &gam, sizeof(double)));
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(gm1_d), &gm1, sizeof(double)));
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(cfl_d), &cfl, sizeof(double)));
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(eps_d), &eps, sizeof(double)));
  }
  hipStream_t st;
  ECHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  ECHECK(hipDeviceSynchronize());

  int nsm = 0;
  ECHECK(hipDeviceGetAttribute(&nsm, hipDeviceAttributeMultiprocessorCount, 0));
  // block sizes of save_soln, adt_calc, res_calc, bres_calc (the defaults, or
  // the autotuner's picks); every kernel's result is independent of them.
  // res_calc's is its owner patch size, fixed by the partition.
  int blk[4] = {256, 1024, kPatch, 128};
  // with `ev`, each launch is bracketed by events ev[2j], ev[2j+1] in launch
  // order: copy, then adt, res, bres of each half step
  auto step_range = [&](int mlo, int mhi, hipEvent_t *ev = nullptr) {
    // save_soln, then twice adt_calc, res_calc, bres_calc, update
    const int c0 = coff[mlo], c1 = coff[mhi];
    long long lo2 = 2LL * c0, hi2 = 2LL * c1;
    int e = 0;
    auto mark = [&]() {
      if (ev) ECHECK(hipEventRecord(ev[e++], st));
    };
    mark();
    k_ens_copy<<<grid_for(hi2 - lo2, blk[0]), blk[0], 0, st>>>((const double2 *)d_q,
                                                              (double2 *)d_qold, lo2, hi2);
    mark();
    for (int k = 0; k < 2; ++k) {
      mark();
      k_ens_adt<<<grid_for(c1 - c0, blk[1]), blk[1], 0, st>>>(d_x, d_q, d_adt, d_cell, c0, c1);
      mark();
      mark();
      if (mblk[mhi] > mblk[mlo])
        k_ens_res_owner<<<mblk[mhi] - mblk[mlo], kPatch, res_lds, st>>>(d_x, d_q, d_adt, d_res,
                                                                        OA, mblk[mlo]);
      mark();
      const int t0 = mesh_btarget_off[mlo], t1 = mesh_btarget_off[mhi];
      mark();
      if (t1 > t0)
        k_ens_bres<<<grid_for(t1 - t0, blk[3]), blk[3], 0, st>>>(
            d_x, d_q, d_adt, d_res, d_bedge, d_bound, d_bmesh, d_qinf, d_boffsets,
            d_bentries, d_btargets, t0, t1);
      mark();
      const int tl0 = mesh_tile_off[mlo], tl1 = mesh_tile_off[mhi];
      const int grid = std::max(1, std::min(tl1 - tl0, 4 * nsm));
      k_ens_update<<<grid, kTile, 0, st>>>(d_qold, d_q, d_res, d_adt, d_tstart, d_tlen,
                                            d_tbase, tl0, tl1, d_psum, d_pmax, d_ploc);
      k_ens_finish<<<mhi - mlo, 256, 0, st>>>(d_psum, d_pmax, d_ploc, d_mtoff, mlo,
                                              d_slots);
    }
  };
  auto history = [&]() {
    k_ens_history<<<1, 256, 0, st>>>(d_slots, hist_words, d_hist, d_count, capacity);
  };

  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<hipEvent_t> ev_a, ev_b;
  std::vector<double> member_ms(M, 0.0);
  int timed_steps = 0;
  if (!o->batched && o->member_timing) {
    ev_a.resize(M);
    ev_b.resize(M);
    for (int m = 0; m < M; ++m) {
      ECHECK(hipEventCreate(&ev_a[m]));
      ECHECK(hipEventCreate(&ev_b[m]));
    }
  }
  // autotuning: step i (i < 5) runs the save_soln, adt_calc and bres_calc
  // kernels with candidate i; the fastest total per kernel over its launches
  // wins
  const int cand[5] = {64, 128, 256, 512, 1024};
  const int ntune = (o->batched && o->autotune) ? std::min(5, o->niter) : 0;
  double tune_ms[4][5] = {};
  hipEvent_t tev[14];
  if (ntune)
    for (auto &e : tev) ECHECK(hipEventCreate(&e));
  double t0 = 0.0;
  for (int it = 1; it <= o->niter; ++it) {
    if (it == o->warmup + 1) {
      ECHECK(hipStreamSynchronize(st));
      res->setup_seconds = now() - t_setup;
      t0 = now();
    }
    if (it <= ntune) {
      blk[0] = blk[1] = blk[3] = cand[it - 1];
      step_range(0, M, tev);
      history();
      ECHECK(hipStreamSynchronize(st));
      // event pairs: copy; then (adt, res, bres) twice
      const int kind[7] = {0, 1, 2, 3, 1, 2, 3};
      for (int j = 0; j < 7; ++j) {
        float ms = 0.f;
        ECHECK(hipEventElapsedTime(&ms, tev[2 * j], tev[2 * j + 1]));
        tune_ms[kind[j]][it - 1] += ms;
      }
      if (it == ntune)
        for (int k : {0, 1, 3}) {
          int best = 0;
          for (int c = 1; c < ntune; ++c)
            if (tune_ms[k][c] < tune_ms[k][best]) best = c;
          blk[k] = cand[best];
        }
    } else if (o->batched) {
      if (o->graphed) {
        if (!gexec) {
          ECHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          step_range(0, M);
          history();
          ECHECK(hipStreamEndCapture(st, &graph));
          ECHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        ECHECK(hipGraphLaunch(gexec, st));
      } else {
        step_range(0, M);
        history();
      }
    } else {
      for (int m = 0; m < M; ++m) {
        if (o->member_timing) ECHECK(hipEventRecord(ev_a[m], st));
        step_range(m, m + 1);
        if (o->member_timing) ECHECK(hipEventRecord(ev_b[m], st));
      }
      history();
      if (o->member_timing && it > o->warmup) {
        ECHECK(hipStreamSynchronize(st));
        for (int m = 0; m < M; ++m) {
          float ms = 0.f;
          ECHECK(hipEventElapsedTime(&ms, ev_a[m], ev_b[m]));
          member_ms[m] += ms;
        }
        ++timed_steps;
      }
    }
    ECHECK(hipGetLastError());
  }
  ECHECK(hipStreamSynchronize(st));
  res->seconds = now() - t0;

  // ---- results ----
  std::vector<double> hist((size_t)hist_words * capacity);
  unsigned long long count = 0;
  ECHECK(hipMemcpy(&count, d_count, sizeof(count), hipMemcpyDeviceToHost));
  ECHECK(hipMemcpy(hist.data(), d_hist, hist.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (count != (unsigned long long)o->niter) {
    printf("op2hip ensemble: %llu history records for %d iterations\n", count, o->niter);
    return 5;
  }
  for (int it = 0; it < o->niter; ++it)
    for (int m = 0; m < M; ++m) {
      const double *h = &hist[(size_t)it * hist_words + 3 * m];
      const int nc = coff[m + 1] - coff[m];
      if (res->rms) res->rms[(size_t)it * M + m] = std::sqrt((0.0 + h[0]) / (double)nc);
      // maxerr/errloc start at 0 (airfoil.cpp:265-267): replaced when exceeded
      if (res->maxerr) res->maxerr[(size_t)it * M + m] = h[1] > 0.0 ? h[1] : 0.0;
      if (res->errloc) res->errloc[(size_t)it * M + m] = h[1] > 0.0 ? (int)h[2] : 0;
    }
  if (res->q_out)
    ECHECK(hipMemcpy(res->q_out, d_q, 4 * (size_t)Nc * sizeof(double), hipMemcpyDeviceToHost));
  if (res->member_step_ms) {
    for (int m = 0; m < M; ++m)
      res->member_step_ms[m] =
          o->batched ? 1e3 * res->seconds / std::max(1, o->niter - o->warmup)
                     : (timed_steps ? member_ms[m] / timed_steps : 0.0);
  }
  for (int k = 0; k < 4; ++k) res->blocks[k] = blk[k];
  if (ntune)
    for (auto e : tev) (void)hipEventDestroy(e);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto e : ev_a) (void)hipEventDestroy(e);
  for (auto e : ev_b) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(st);
  for (void *p : keep) (void)hipFree(p);
  return 0;
}
