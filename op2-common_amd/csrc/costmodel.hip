// MLP cost model (include/op2amd_costmodel.h): forward, backward and Adam as
// HIP kernels.  Every product of the forward and backward passes (the dense
// layers, the weight and bias gradients, the back-propagated deltas) runs on
// the matrix cores (v_mfma_f32_16x16x4f32, FP32 in and out); the weight
// gradients, which reduce over the samples, are split over the samples in
// fixed chunks whose partials are summed in chunk order, so the gradients are
// deterministic.
#include "op2amd_costmodel.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#define MCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip mlp: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__, \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

struct op2amd_mlp {
  int nin, H, np;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;  // device
  // device work buffers, grown on demand
  int cap = 0;
  float *X = nullptr, *T = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr,
        *A2 = nullptr, *Y = nullptr, *DY = nullptr, *D2 = nullptr, *D1 = nullptr,
        *PK = nullptr, *loss = nullptr;  // PK: split partials of the gradients
  int adam_t = 0;
  float wd = 0.f;  // decoupled weight decay
  hipStream_t st = nullptr;
  // parameter offsets
  int oW1, ob1, oW2, ob2, oW3, ob3;
};

namespace {

int blocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kSplit = 256;  // samples per split of a weight-gradient product

int splits_for(int n) { return (n + kSplit - 1) / kSplit; }

// C(m, n) = sum over k of a(m, k) b(k, n), k in split blockIdx.z's range
// [z kchunk, (z + 1) kchunk) of [0, K), with
//   a(m, k) = A[m sam + k sak] for m < ma, 1 for ma <= m < M (a row of ones
//             gives the bias gradient next to the weight gradient),
//   b(k, n) = B[k sbk + n sbn],
// on the matrix cores: each wave one 16x16 tile of C, four k per
// v_mfma_f32_16x16x4f32 (A operand: lane l holds a(l % 16, l / 16); B
// operand: b(l / 16, l % 16); result: lane l, element v holds
// C(4 (l / 16) + v, l % 16)).  Operands outside the matrices are zeros.
// Epilogue: EP 0 stores split z's partial at C + z M N; EP 1 stores
// Z = C + bias[n] and, when R is given, R = relu(Z); EP 2 stores C where
// mask(m, n) > 0, else 0.  C, R and mask are row-major M x N.
template <int EP>
__global__ void __launch_bounds__(256)
    k_gemm(const float *__restrict A, int sam, int sak, int ma, const float *__restrict B,
           int sbk, int sbn, float *__restrict C, float *__restrict R,
           const float *__restrict aux, int M, int N, int K, int kchunk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i0 = (blockIdx.x * 4 + wave) * 16, j0 = blockIdx.y * 16;
  if (i0 >= M) return;  // the whole wave: no barriers below
  const int k_lo = blockIdx.z * kchunk, k_hi = min(K, k_lo + kchunk);
  const int ia = i0 + (lane & 15), kk = lane >> 4, jb = j0 + (lane & 15);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = k_lo; k0 < k_hi; k0 += 4) {
    const int k = k0 + kk;
    float a = 0.f, b = 0.f;
    if (k < k_hi) {
      if (ia < ma)
        a = A[(size_t)ia * sam + (size_t)k * sak];
      else if (ia < M)
        a = 1.f;
      if (jb < N) b = B[(size_t)k * sbk + (size_t)jb * sbn];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
  const int j = j0 + (lane & 15);
  if (j >= N) return;
  for (int v = 0; v < 4; ++v) {
    const int i = i0 + 4 * (lane >> 4) + v;
    if (i >= M) break;
    const size_t o = (size_t)i * N + j;
    if (EP == 0) {
      C[(size_t)blockIdx.z * M * N + o] = acc[v];
    } else if (EP == 1) {
      const float z = acc[v] + aux[j];
      C[o] = z;
      if (R) R[o] = z > 0.f ? z : 0.f;
    } else {
      C[o] = aux[o] > 0.f ? acc[v] : 0.f;
    }
  }
}

// out[i] = sum over splits z, in order, of P[z MN + i]
__global__ void k_split_sum(const float *__restrict P, int splits, int MN,
                            float *__restrict out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float s = P[i];
  for (int z = 1; z < splits; ++z) s += P[(size_t)z * MN + i];
  out[i] = s;
}

dim3 gemm_grid(int M, int N, int splits) {
  return dim3((unsigned)(((M + 15) / 16 + 3) / 4), (unsigned)((N + 15) / 16), (unsigned)splits);
}

// DY[n] = 2 (Y[n] - T[n]) / N
__global__ void k_dloss(const float *__restrict Y, const float *__restrict T,
                        float *__restrict DY, int N) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) DY[n] = 2.f * (Y[n] - T[n]) / (float)N;
}

// loss = mean (Y - T)^2, one workgroup, lanes in order then a tree
__global__ void k_loss(const float *__restrict Y, const float *__restrict T,
                       float *__restrict loss, int N) {
  __shared__ float sh[256];
  float s = 0.f;
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    float d = Y[n] - T[n];
    s += d * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = sh[0] / (float)N;
}

// Adam with decoupled weight decay `wd` (AdamW; wd = 0 is plain Adam).
__global__ void k_adam(float *__restrict P, const float *__restrict G,
                       float *__restrict M, float *__restrict V, int np, float lr,
                       float c1, float c2, float wd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  float g = G[i];
  float m = 0.9f * M[i] + 0.1f * g;
  float v = 0.999f * V[i] + 0.001f * g * g;
  M[i] = m;
  V[i] = v;
  float p = P[i];
  P[i] = p - lr * (m / c1) / (sqrtf(v / c2) + 1e-8f) - lr * wd * p;
}

void ensure(op2amd_mlp *m, int n) {
  if (n <= m->cap) return;
  float **bufs[] = {&m->X, &m->T, &m->Z1, &m->A1, &m->Z2, &m->A2, &m->Y, &m->DY, &m->D2, &m->D1,
                    &m->PK};
  for (float **b : bufs)
    if (*b) MCHECK(hipFree(*b));
  const size_t H = m->H;
  MCHECK(hipMalloc(&m->X, sizeof(float) * n * m->nin));
  MCHECK(hipMalloc(&m->T, sizeof(float) * n));
  MCHECK(hipMalloc(&m->Z1, sizeof(float) * n * H));
  MCHECK(hipMalloc(&m->A1, sizeof(float) * n * H));
  MCHECK(hipMalloc(&m->Z2, sizeof(float) * n * H));
  MCHECK(hipMalloc(&m->A2, sizeof(float) * n * H));
  MCHECK(hipMalloc(&m->Y, sizeof(float) * n));
  MCHECK(hipMalloc(&m->DY, sizeof(float) * n));
  MCHECK(hipMalloc(&m->D2, sizeof(float) * n * H));
  MCHECK(hipMalloc(&m->D1, sizeof(float) * n * H));
  // split partials of the largest gradient product: (max(nin, H) + 1) x H
  MCHECK(hipMalloc(&m->PK, sizeof(float) * splits_for(n) * (std::max(m->nin, m->H) + 1) * H));
  m->cap = n;
}

// Z = A W + b (A: n x K, W: K x Mo), R = relu(Z) when R is given
void dense(op2amd_mlp *m, const float *A, int n, int K, const float *W, const float *b,
           int Mo, float *Z, float *R) {
  k_gemm<1><<<gemm_grid(n, Mo, 1), 256, 0, m->st>>>(A, K, 1, n, W, Mo, 1, Z, R, b, n, Mo, K, K);
}

// [dW; db] = [A^T; 1] D over the n samples (A: n x K, D: n x Mo) into the
// (K + 1) x Mo block G (W's gradient, then b's, as the parameters are laid out)
void wgrad(op2amd_mlp *m, const float *A, int n, int K, const float *D, int Mo, float *G) {
  const int sp = splits_for(n), M = K + 1;
  k_gemm<0><<<gemm_grid(M, Mo, sp), 256, 0, m->st>>>(A, 1, K, K, D, Mo, 1, m->PK, nullptr,
                                                       nullptr, M, Mo, n, kSplit);
  k_split_sum<<<blocks((long long)M * Mo, 256), 256, 0, m->st>>>(m->PK, sp, M * Mo, G);
}

// Din = (D W^T) where Z > 0, else 0 (D: n x Mo, W: K x Mo, Z and Din: n x K)
void back(op2amd_mlp *m, const float *D, int n, int Mo, const float *W, int K, const float *Z,
          float *Din) {
  k_gemm<2><<<gemm_grid(n, K, 1), 256, 0, m->st>>>(D, Mo, 1, n, W, 1, Mo, Din, nullptr, Z, n, K,
                                                     Mo, Mo);
}

void forward(op2amd_mlp *m, int n) {
  const int H = m->H;
  float *P = m->P;
  dense(m, m->X, n, m->nin, P + m->oW1, P + m->ob1, H, m->Z1, m->A1);
  dense(m, m->A1, n, H, P + m->oW2, P + m->ob2, H, m->Z2, m->A2);
  dense(m, m->A2, n, H, P + m->oW3, P + m->ob3, 1, m->Y, nullptr);
}

void backward(op2amd_mlp *m, int n) {
  const int H = m->H;
  float *P = m->P, *G = m->G;
  k_dloss<<<blocks(n, 256), 256, 0, m->st>>>(m->Y, m->T, m->DY, n);
  wgrad(m, m->A2, n, H, m->DY, 1, G + m->oW3);
  back(m, m->DY, n, 1, P + m->oW3, H, m->Z2, m->D2);
  wgrad(m, m->A1, n, H, m->D2, H, G + m->oW2);
  back(m, m->D2, n, H, P + m->oW2, H, m->Z1, m->D1);
  wgrad(m, m->X, n, m->nin, m->D1, H, G + m->oW1);
}

void upload(op2amd_mlp *m, const float *x, const float *t, int n) {
  ensure(m, n);
  MCHECK(hipMemcpyAsync(m->X, x, sizeof(float) * n * m->nin, hipMemcpyHostToDevice, m->st));
  if (t) MCHECK(hipMemcpyAsync(m->T, t, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
}

}  // namespace

extern "C" op2amd_mlp *op2amd_mlp_create(int nin, int hidden, uint64_t seed) {
  int ndev = 0;
  if (nin < 1 || hidden < 1 || hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return nullptr;
  op2amd_mlp *m = new op2amd_mlp();
  m->nin = nin;
  m->H = hidden;
  m->oW1 = 0;
  m->ob1 = m->oW1 + nin * hidden;
  m->oW2 = m->ob1 + hidden;
  m->ob2 = m->oW2 + hidden * hidden;
  m->oW3 = m->ob2 + hidden;
  m->ob3 = m->oW3 + hidden;
  m->np = m->ob3 + 1;
  MCHECK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
  MCHECK(hipMalloc(&m->P, sizeof(float) * m->np));
  MCHECK(hipMalloc(&m->G, sizeof(float) * m->np));
  MCHECK(hipMalloc(&m->M, sizeof(float) * m->np));
  MCHECK(hipMalloc(&m->V, sizeof(float) * m->np));
  MCHECK(hipMalloc(&m->loss, sizeof(float)));
  MCHECK(hipMemset(m->M, 0, sizeof(float) * m->np));
  MCHECK(hipMemset(m->V, 0, sizeof(float) * m->np));
  MCHECK(hipDeviceSynchronize());  // null-stream work done before m->st uses it
  // He initialisation: normal(0, sqrt(2/fan_in)) by Box-Muller on SplitMix64
  std::vector<float> p(m->np, 0.f);
  uint64_t s = seed;
  auto uni = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
  };
  auto normal = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(2 * M_PI * uni()); };
  for (int i = 0; i < nin * hidden; ++i) p[m->oW1 + i] = (float)(normal() * std::sqrt(2.0 / nin));
  for (int i = 0; i < hidden * hidden; ++i) p[m->oW2 + i] = (float)(normal() * std::sqrt(2.0 / hidden));
  for (int i = 0; i < hidden; ++i) p[m->oW3 + i] = (float)(normal() * std::sqrt(1.0 / hidden));
  MCHECK(hipMemcpy(m->P, p.data(), sizeof(float) * m->np, hipMemcpyHostToDevice));
  return m;
}

extern "C" void op2amd_mlp_destroy(op2amd_mlp *m) {
  if (!m) return;
  float *bufs[] = {m->P, m->G, m->M, m->V, m->loss, m->X, m->T, m->Z1, m->A1,
                   m->Z2, m->A2, m->Y, m->DY, m->D2, m->D1, m->PK};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
}

extern "C" int op2amd_mlp_num_params(const op2amd_mlp *m) { return m->np; }

extern "C" int op2amd_mlp_get_params(const op2amd_mlp *m, float *out) {
  MCHECK(hipStreamSynchronize(m->st));
  MCHECK(hipMemcpy(out, m->P, sizeof(float) * m->np, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int op2amd_mlp_set_params(op2amd_mlp *m, const float *in) {
  MCHECK(hipStreamSynchronize(m->st));
  MCHECK(hipMemcpy(m->P, in, sizeof(float) * m->np, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int op2amd_mlp_forward(op2amd_mlp *m, const float *x, int n, float *y) {
  if (n < 1) return 1;
  upload(m, x, nullptr, n);
  forward(m, n);
  MCHECK(hipMemcpyAsync(y, m->Y, sizeof(float) * n, hipMemcpyDeviceToHost, m->st));
  MCHECK(hipStreamSynchronize(m->st));
  return 0;
}

extern "C" int op2amd_mlp_loss_grad(op2amd_mlp *m, const float *x, const float *t, int n,
                                    float *loss, float *grad) {
  if (n < 1) return 1;
  upload(m, x, t, n);
  forward(m, n);
  backward(m, n);
  k_loss<<<1, 256, 0, m->st>>>(m->Y, m->T, m->loss, n);
  MCHECK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
  MCHECK(hipMemcpyAsync(grad, m->G, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->st));
  MCHECK(hipStreamSynchronize(m->st));
  return 0;
}

extern "C" int op2amd_mlp_set_weight_decay(op2amd_mlp *m, float wd) {
  if (!m || wd < 0.f) return 1;
  m->wd = wd;
  return 0;
}

extern "C" int op2amd_mlp_train(op2amd_mlp *m, const float *x, const float *t, int n,
                                int epochs, float lr, float *final_loss) {
  if (n < 1) return 1;
  upload(m, x, t, n);
  for (int e = 0; e < epochs; ++e) {
    forward(m, n);
    backward(m, n);
    ++m->adam_t;
    float c1 = 1.f - std::pow(0.9f, (float)m->adam_t);
    float c2 = 1.f - std::pow(0.999f, (float)m->adam_t);
    k_adam<<<blocks(m->np, 256), 256, 0, m->st>>>(m->P, m->G, m->M, m->V, m->np, lr, c1, c2,
                                                  m->wd);
  }
  forward(m, n);
  k_loss<<<1, 256, 0, m->st>>>(m->Y, m->T, m->loss, n);
  float l = 0.f;
  MCHECK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
  MCHECK(hipStreamSynchronize(m->st));
  if (final_loss) *final_loss = l;
  return 0;
}
