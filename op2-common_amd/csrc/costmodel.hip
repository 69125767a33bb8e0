// MLP cost model (include/op2amd_costmodel.h): forward, backward and Adam as
// HIP kernels.  Sizes are small (a few thousand samples, 64 hidden units), so
// every kernel assigns one thread per output element and reduces over samples
// sequentially — deterministic gradients.
#include "op2amd_costmodel.h"

#include <hip/hip_runtime.h>

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
        *loss = nullptr;
  int adam_t = 0;
  float wd = 0.f;  // decoupled weight decay
  hipStream_t st = nullptr;
  // parameter offsets
  int oW1, ob1, oW2, ob2, oW3, ob3;
};

namespace {

int blocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// Z[n][m] = b[m] + sum_k X[n][k] W[k][m];  A = relu(Z) (A may be null)
__global__ void k_dense(const float *__restrict X, const float *__restrict W,
                        const float *__restrict b, float *__restrict Z,
                        float *__restrict A, int N, int K, int Mo) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)N * Mo) return;
  int n = (int)(i / Mo), m = (int)(i % Mo);
  float z = b[m];
  for (int k = 0; k < K; ++k) z += X[(size_t)n * K + k] * W[(size_t)k * Mo + m];
  Z[i] = z;
  if (A) A[i] = z > 0.f ? z : 0.f;
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

// dW[k][m] = sum_n A[n][k] D[n][m];  db[m] = sum_n D[n][m]  (db when k == K)
__global__ void k_wgrad(const float *__restrict A, const float *__restrict D,
                        float *__restrict dW, float *__restrict db, int N, int K,
                        int Mo) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (K + 1) * Mo) return;
  int k = i / Mo, m = i % Mo;
  float s = 0.f;
  if (k < K) {
    for (int n = 0; n < N; ++n) s += A[(size_t)n * K + k] * D[(size_t)n * Mo + m];
    dW[(size_t)k * Mo + m] = s;
  } else {
    for (int n = 0; n < N; ++n) s += D[(size_t)n * Mo + m];
    db[m] = s;
  }
}

// Din[n][k] = (sum_m D[n][m] W[k][m]) * (Z[n][k] > 0)
__global__ void k_back(const float *__restrict D, const float *__restrict W,
                       const float *__restrict Z, float *__restrict Din, int N, int K,
                       int Mo) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)N * K) return;
  int n = (int)(i / K), k = (int)(i % K);
  float s = 0.f;
  for (int m = 0; m < Mo; ++m) s += D[(size_t)n * Mo + m] * W[(size_t)k * Mo + m];
  Din[i] = Z[i] > 0.f ? s : 0.f;
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
  float **bufs[] = {&m->X, &m->T, &m->Z1, &m->A1, &m->Z2, &m->A2, &m->Y, &m->DY, &m->D2, &m->D1};
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
  m->cap = n;
}

void forward(op2amd_mlp *m, int n) {
  const int H = m->H;
  float *P = m->P;
  k_dense<<<blocks((long long)n * H, 256), 256, 0, m->st>>>(m->X, P + m->oW1, P + m->ob1, m->Z1, m->A1, n, m->nin, H);
  k_dense<<<blocks((long long)n * H, 256), 256, 0, m->st>>>(m->A1, P + m->oW2, P + m->ob2, m->Z2, m->A2, n, H, H);
  k_dense<<<blocks(n, 256), 256, 0, m->st>>>(m->A2, P + m->oW3, P + m->ob3, m->Y, nullptr, n, H, 1);
}

void backward(op2amd_mlp *m, int n) {
  const int H = m->H;
  float *P = m->P, *G = m->G;
  k_dloss<<<blocks(n, 256), 256, 0, m->st>>>(m->Y, m->T, m->DY, n);
  k_wgrad<<<blocks((long long)(H + 1) * 1, 256), 256, 0, m->st>>>(m->A2, m->DY, G + m->oW3, G + m->ob3, n, H, 1);
  k_back<<<blocks((long long)n * H, 256), 256, 0, m->st>>>(m->DY, P + m->oW3, m->Z2, m->D2, n, H, 1);
  k_wgrad<<<blocks((long long)(H + 1) * H, 256), 256, 0, m->st>>>(m->A1, m->D2, G + m->oW2, G + m->ob2, n, H, H);
  k_back<<<blocks((long long)n * H, 256), 256, 0, m->st>>>(m->D2, P + m->oW2, m->Z1, m->D1, n, H, H);
  k_wgrad<<<blocks((long long)(m->nin + 1) * H, 256), 256, 0, m->st>>>(m->X, m->D1, G + m->oW1, G + m->ob1, n, m->nin, H);
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
                   m->Z2, m->A2, m->Y, m->DY, m->D2, m->D1};
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
