// Learned per-sphere eigenstrain predictor (include/ibem3d.h, ibem3d_gnn_*): a message-passing MLP
// over every sphere's kNN = 16 nearest neighbours, trained and evaluated by the kernels here (no
// library). No counterpart in the reference; it supplies the initial guess of the block-Jacobi GMRES.
//
// Graph (bit-exact): centres quantised to integers, c_q = rint(c * 2^kQBits), squared distances in
// 64-bit integers, neighbours ordered by (distance, index); spheres with fewer than kNN others get
// padding slots (mask 0).
// Features of the edge i -> j (kFe), with d = (c_j - c_i) / r_i and s = r_j / r_i: the unit direction
// d / |d| (3), s, the decay of a sphere's disturbance strain at the distance, (s / |d|)^3, and of its
// gradient, s^3 / |d|^4; slot 6 holds the mask, 7 is padding. Node features (kFn): log(E1 / E0),
// nu1, min(distance to a cube face / r_i, 10) / 10; slot 3 is padding.
// Network (FP32 parameters, weights stored transposed, Wt[k][n]):
//   z1_ij = relu(W1 e_ij + b1)                  (128)
//   a_i   = sum_j m_ij z1_ij / max(1, sum_j m_ij)
//   z2_i  = relu(W2 [a_i; f_i] + b2)            (128)
//   z3_i  = relu(W3 z2_i + b3)                  (128)
//   y_i   = W4 z3_i + b4                        (360 = 60 coefficients x 6 unit macro strains)
// Loss: mean of (y - target)^2 over the minibatch's spheres and 360 outputs. Training: Adam.
// Every reduction (weight gradients, loss) runs in a fixed order: results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ibem3d.h"

namespace {

constexpr int kNN = 16;
constexpr int kFe = 6, kFeS = 8;  // edge features, stride
constexpr int kFn = 3, kFnS = 4;  // node features, stride
constexpr int kH = 128;
constexpr int kOut = 360;
constexpr int kIn2 = kH + kFn;
constexpr int kQBits = 28;

// parameter layout (floats): W1t [kFe][kH], b1, W2t [kIn2][kH], b2, W3t [kH][kH], b3, W4t [kH][kOut], b4
constexpr long kOffW1 = 0, kOffB1 = kOffW1 + kFe * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kIn2 * kH,
               kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kH, kOffW4 = kOffB3 + kH,
               kOffB4 = kOffW4 + kH * kOut, kParams = kOffB4 + kOut;

std::string g_gnn_error;

#define GCHK(call)                                                                       \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess) {                                                              \
      g_gnn_error = std::string(#call) + ": " + hipGetErrorString(_e);                   \
      return IBEM3D_EHIP;                                                                \
    }                                                                                    \
  } while (0)

// ---------------------------------------------------------------------------- graph + features ---
__global__ __launch_bounds__(256) void k_knn(const double* __restrict__ cen, int num, int* __restrict__ nbr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  const double s = (double)(1 << kQBits);
  const long long xi = llrint(cen[3 * i] * s), yi = llrint(cen[3 * i + 1] * s), zi = llrint(cen[3 * i + 2] * s);
  unsigned long long bd[kNN];
  int bj[kNN];
  for (int q = 0; q < kNN; ++q) {
    bd[q] = ~0ull;
    bj[q] = -1;
  }
  for (int j = 0; j < num; ++j) {
    if (j == i) continue;
    const long long dx = llrint(cen[3 * j] * s) - xi, dy = llrint(cen[3 * j + 1] * s) - yi,
                    dz = llrint(cen[3 * j + 2] * s) - zi;
    const unsigned long long d = (unsigned long long)(dx * dx) + (unsigned long long)(dy * dy) +
                                 (unsigned long long)(dz * dz);
    // insert (d, j) into the sorted list: ties keep the smaller index first (j grows)
    if (d >= bd[kNN - 1]) continue;
    int q = kNN - 1;
    while (q > 0 && bd[q - 1] > d) {
      bd[q] = bd[q - 1];
      bj[q] = bj[q - 1];
      --q;
    }
    bd[q] = d;
    bj[q] = j;
  }
  for (int q = 0; q < kNN; ++q) nbr[(long)i * kNN + q] = bj[q];
}

__global__ __launch_bounds__(256) void k_features(const double* __restrict__ cen, const double* __restrict__ rad,
                                                  int num, const int* __restrict__ nbr, float le, float nu1,
                                                  float* __restrict__ edge, float* __restrict__ node) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= num) return;
  const double ri = rad[i];
  double wall = 1e300;
  for (int c = 0; c < 3; ++c) wall = fmin(wall, fmin(cen[3 * i + c], 1.0 - cen[3 * i + c]));
  float* f = node + (long)i * kFnS;
  f[0] = le;
  f[1] = nu1;
  f[2] = (float)(fmin(wall / ri, 10.0) / 10.0);
  f[3] = 0.0f;
  for (int q = 0; q < kNN; ++q) {
    float* e = edge + ((long)i * kNN + q) * kFeS;
    const int j = nbr[(long)i * kNN + q];
    if (j < 0) {
      for (int u = 0; u < kFeS; ++u) e[u] = 0.0f;
      continue;
    }
    const double d0 = (cen[3 * j] - cen[3 * i]) / ri, d1 = (cen[3 * j + 1] - cen[3 * i + 1]) / ri,
                 d2 = (cen[3 * j + 2] - cen[3 * i + 2]) / ri;
    const double dn = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double sj = rad[j] / ri, q3 = sj * sj * sj / (dn * dn * dn);
    e[0] = (float)(d0 / dn);
    e[1] = (float)(d1 / dn);
    e[2] = (float)(d2 / dn);
    e[3] = (float)sj;
    e[4] = (float)q3;
    e[5] = (float)(q3 / dn);
    e[6] = 1.0f;
    e[7] = 0.0f;
  }
}

// ----------------------------------------------------------------------------- dense layers -----
// Y[m][n] = act(sum_k X[m][k] Wt[k][n] + b[n]); X rows of stride ldx. One thread per output; a
// workgroup takes 2 rows x 128 columns, the two rows of X staged in LDS.
__global__ __launch_bounds__(256) void k_dense_fwd(const float* __restrict__ X, long M, int K, int ldx,
                                                   const float* __restrict__ Wt, const float* __restrict__ b,
                                                   int N, float* __restrict__ Y, int relu) {
  __shared__ float sx[2][kIn2 + kH];
  const long m0 = (long)blockIdx.y * 2;
  const int n = blockIdx.x * 128 + (threadIdx.x & 127), r = threadIdx.x >> 7;
  for (int u = threadIdx.x; u < 2 * K; u += 256) {
    const int rr = u / K, k = u % K;
    sx[rr][k] = m0 + rr < M ? X[(m0 + rr) * ldx + k] : 0.0f;
  }
  __syncthreads();
  const long m = m0 + r;
  if (m >= M || n >= N) return;
  float acc = b[n];
  for (int k = 0; k < K; ++k) acc += sx[r][k] * Wt[(long)k * N + n];
  Y[m * N + n] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dX[m][k] = sum_n dY[m][n] Wt[k][n] (then times the relu mask of X's producer when given:
// mask[m][k] > 0). One thread per (m, k); dY's row staged in LDS.
__global__ __launch_bounds__(256) void k_dense_bwd_x(const float* __restrict__ dY, long M, int N,
                                                     const float* __restrict__ Wt, int K,
                                                     const float* __restrict__ act, float* __restrict__ dX) {
  __shared__ float sd[kOut];
  const long m = blockIdx.x;
  for (int u = threadIdx.x; u < N; u += 256) sd[u] = dY[m * N + u];
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += 256) {
    const float* w = Wt + (long)k * N;
    float acc = 0.0f;
    for (int n = 0; n < N; ++n) acc += sd[n] * w[n];
    if (act && !(act[m * K + k] > 0.0f)) acc = 0.0f;
    dX[m * K + k] = acc;
  }
}

// Partial weight gradients over row chunks: part[c][k][n] = sum_{m in chunk c} X[m][k] dY[m][n]
// (k < K), part[c][K][n] = sum dY[m][n] (the bias); one thread per (k, n), rows in order.
constexpr int kChunk = 512;
__global__ __launch_bounds__(256) void k_dense_bwd_w(const float* __restrict__ X, int ldx,
                                                     const float* __restrict__ dY, long M, int K, int N,
                                                     float* __restrict__ part) {
  const long c = blockIdx.y;
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= (long)(K + 1) * N) return;
  const int k = (int)(u / N), n = (int)(u % N);
  const long m0 = c * kChunk, m1 = min(M, m0 + kChunk);
  float acc = 0.0f;
  for (long m = m0; m < m1; ++m) acc += (k < K ? X[m * ldx + k] : 1.0f) * dY[m * N + n];
  part[c * (long)(K + 1) * N + u] = acc;
}
// g[u] = sum over chunks of part (u < (K + 1) N: weights then bias), in chunk order
__global__ __launch_bounds__(256) void k_sum_chunks(const float* __restrict__ part, long chunks, long len,
                                                    float* __restrict__ gw, float* __restrict__ gb, int KN) {
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= len) return;
  float acc = 0.0f;
  for (long c = 0; c < chunks; ++c) acc += part[c * len + u];
  if (u < KN)
    gw[u] = acc;
  else
    gb[u - KN] = acc;
}

// masked mean over the kNN edges of each sphere; ldo = row stride of the output (it feeds W2 with
// the node features appended)
__global__ __launch_bounds__(128) void k_aggregate(const float* __restrict__ z1, const float* __restrict__ edge,
                                                   const float* __restrict__ node, long S,
                                                   float* __restrict__ in2) {
  const long i = blockIdx.x;
  const int h = threadIdx.x;
  float acc = 0.0f, cnt = 0.0f;
  for (int q = 0; q < kNN; ++q) {
    const float m = edge[(i * kNN + q) * kFeS + 6];
    acc += m * z1[(i * kNN + q) * kH + h];
    cnt += m;
  }
  in2[i * kIn2 + h] = acc / fmaxf(cnt, 1.0f);
  if (h < kFn) in2[i * kIn2 + kH + h] = node[i * kFnS + h];
}
// dz1_ij = m_ij da_i / max(1, sum m), times relu'(z1)
__global__ __launch_bounds__(128) void k_aggregate_bwd(const float* __restrict__ din2, const float* __restrict__ edge,
                                                       const float* __restrict__ z1, long S,
                                                       float* __restrict__ dz1) {
  const long i = blockIdx.x;
  const int h = threadIdx.x;
  float cnt = 0.0f;
  for (int q = 0; q < kNN; ++q) cnt += edge[(i * kNN + q) * kFeS + 6];
  const float d = din2[i * kIn2 + h] / fmaxf(cnt, 1.0f);
  for (int q = 0; q < kNN; ++q) {
    const long e = i * kNN + q;
    const float m = edge[e * kFeS + 6];
    dz1[e * kH + h] = (m > 0.0f && z1[e * kH + h] > 0.0f) ? d : 0.0f;
  }
}

// loss: dY = 2 (y - t) / (S kOut), part[b] = block sums of (y - t)^2 (fixed order)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t, long len,
                                             float* __restrict__ dy, float* __restrict__ part) {
  __shared__ float red[256];
  float acc = 0.0f;
  const float sc = 2.0f / (float)len;
  for (long u = (long)blockIdx.x * 256 + threadIdx.x; u < len; u += (long)gridDim.x * 256) {
    const float d = y[u] - t[u];
    dy[u] = sc * d;
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                              float c1, float c2) {
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= n) return;
  const float b1 = 0.9f, b2 = 0.999f;
  m[u] = b1 * m[u] + (1.0f - b1) * g[u];
  v[u] = b2 * v[u] + (1.0f - b2) * g[u] * g[u];
  p[u] -= lr * (m[u] / c1) / (sqrtf(v[u] / c2) + 1e-8f);
}

// gather a minibatch: rows idx[0..B) of the dataset
__global__ __launch_bounds__(256) void k_gather(const int* __restrict__ idx, long B, const float* __restrict__ src,
                                                long width, float* __restrict__ dst) {
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= B * width) return;
  const long r = u / width, c = u % width;
  dst[u] = src[(long)idx[r] * width + c];
}

// ------------------------------------------------------------------- BF16 inference on MFMA ---
// Y = act(X Wt + b) with X (M x K, stride ldx) and Wt (K x N) rounded to BF16 and FP32 accumulation
// in v_mfma_f32_16x16x32_bf16: one wavefront per 16 x 16 output tile, K in steps of 32 (zero beyond
// K). Fragment layout: lane l holds A[l % 16][8 (l / 16) + q] and B[8 (l / 16) + q][l % 16],
// q < 8; accumulator register v of lane l is D[4 (l / 16) + v][l % 16].
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_dense_bf16(const float* __restrict__ X, long M, int K, int ldx,
                                                    const float* __restrict__ Wt, const float* __restrict__ b,
                                                    int N, float* __restrict__ Y, int relu) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long tiles_n = (N + 15) / 16;
  const long tile = (long)blockIdx.x * 4 + wave;
  const long tm = tile / tiles_n, tn = tile % tiles_n;
  if (tm * 16 >= M) return;
  const long m0 = tm * 16, n0 = tn * 16;
  const int lr = lane & 15, lk = lane >> 4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a, w;
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + 8 * lk + q;
      const long m = m0 + lr, n = n0 + lr;
      a[q] = (__bf16)(k < K && m < M ? X[m * ldx + k] : 0.0f);
      w[q] = (__bf16)(k < K && n < N ? Wt[(long)k * N + n] : 0.0f);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc, 0, 0, 0);
  }
  const long n = n0 + lr;
  if (n >= N) return;
  for (int v = 0; v < 4; ++v) {
    const long m = m0 + 4 * lk + v;
    if (m < M) {
      const float y = acc[v] + b[n];
      Y[m * N + n] = relu ? fmaxf(y, 0.0f) : y;
    }
  }
}

unsigned grid1(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct ibem3d_gnn {
  int dev = 0;
  hipStream_t st = nullptr;
  float *p = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;  // parameters, gradient, Adam moments
  long adam_t = 0;
  // dataset (device)
  float *d_edge = nullptr, *d_node = nullptr, *d_tgt = nullptr;
  long S = 0;
  // minibatch buffers (capacity cap spheres)
  long cap = 0;
  float *b_edge = nullptr, *b_node = nullptr, *b_tgt = nullptr, *z1 = nullptr, *in2 = nullptr, *z2 = nullptr,
        *z3 = nullptr, *y = nullptr, *dy = nullptr, *dz3 = nullptr, *dz2 = nullptr, *din2 = nullptr,
        *dz1 = nullptr, *part = nullptr, *lpart = nullptr;
  int* b_idx = nullptr;
};

namespace {

int ensure_batch(ibem3d_gnn* G, long B) {
  if (B <= G->cap) return 0;
  void* bufs[] = {G->b_edge, G->b_node, G->b_tgt, G->z1, G->in2, G->z2, G->z3, G->y, G->dy, G->dz3,
                  G->dz2, G->din2, G->dz1, G->part, G->lpart, G->b_idx};
  for (void* q : bufs) hipFree(q);
  const long E = B * kNN;
  const long chunks = (E + kChunk - 1) / kChunk;
  const long partf = chunks * (long)(kIn2 + 1) * kOut;  // the largest (K + 1) N over the layers
  GCHK(hipMalloc((void**)&G->b_edge, sizeof(float) * E * kFeS));
  GCHK(hipMalloc((void**)&G->b_node, sizeof(float) * B * kFnS));
  GCHK(hipMalloc((void**)&G->b_tgt, sizeof(float) * B * kOut));
  GCHK(hipMalloc((void**)&G->z1, sizeof(float) * E * kH));
  GCHK(hipMalloc((void**)&G->in2, sizeof(float) * B * kIn2));
  GCHK(hipMalloc((void**)&G->z2, sizeof(float) * B * kH));
  GCHK(hipMalloc((void**)&G->z3, sizeof(float) * B * kH));
  GCHK(hipMalloc((void**)&G->y, sizeof(float) * B * kOut));
  GCHK(hipMalloc((void**)&G->dy, sizeof(float) * B * kOut));
  GCHK(hipMalloc((void**)&G->dz3, sizeof(float) * B * kH));
  GCHK(hipMalloc((void**)&G->dz2, sizeof(float) * B * kH));
  GCHK(hipMalloc((void**)&G->din2, sizeof(float) * B * kIn2));
  GCHK(hipMalloc((void**)&G->dz1, sizeof(float) * E * kH));
  GCHK(hipMalloc((void**)&G->part, sizeof(float) * partf));
  GCHK(hipMalloc((void**)&G->lpart, sizeof(float) * 1024));
  GCHK(hipMalloc((void**)&G->b_idx, sizeof(int) * B));
  G->cap = B;
  return 0;
}

void dense_fwd(hipStream_t st, const float* X, long M, int K, int ldx, const float* Wt, const float* b, int N,
               float* Y, bool relu, bool bf16) {
  if (bf16) {
    const long tiles = ((M + 15) / 16) * ((N + 15) / 16);
    hipLaunchKernelGGL(k_dense_bf16, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, st, X, M, K, ldx, Wt, b, N,
                       Y, relu ? 1 : 0);
  } else {
    hipLaunchKernelGGL(k_dense_fwd, dim3((unsigned)((N + 127) / 128), (unsigned)((M + 1) / 2)), dim3(256), 0, st,
                       X, M, K, ldx, Wt, b, N, Y, relu ? 1 : 0);
  }
}

// forward over B spheres whose edges / nodes are at e, f, with the activations in z1 (B x 16 x kH),
// in2 (B x kIn2), z2, z3 (B x kH) and the output in y (B x kOut); every activation element is written
// before it is read
void forward_into(hipStream_t st, const float* P, const float* e, const float* f, long B, bool bf16, float* z1,
                  float* in2, float* z2, float* z3, float* y) {
  const long E = B * kNN;
  dense_fwd(st, e, E, kFe, kFeS, P + kOffW1, P + kOffB1, kH, z1, true, bf16);
  hipLaunchKernelGGL(k_aggregate, dim3((unsigned)B), dim3(128), 0, st, z1, e, f, B, in2);
  dense_fwd(st, in2, B, kIn2, kIn2, P + kOffW2, P + kOffB2, kH, z2, true, bf16);
  dense_fwd(st, z2, B, kH, kH, P + kOffW3, P + kOffB3, kH, z3, true, bf16);
  dense_fwd(st, z3, B, kH, kH, P + kOffW4, P + kOffB4, kOut, y, false, bf16);
}

void forward(ibem3d_gnn* G, const float* e, const float* f, long B, bool bf16) {
  forward_into(G->st, G->p, e, f, B, bf16, G->z1, G->in2, G->z2, G->z3, G->y);
}

// ---- the prediction that allocates nothing: its activations live in a file-scope device array ----
// 16 MiB per device, held until the process ends; every call on a device shares it (calls that overlap
// on one device race). Spheres pass through it in chunks of kScratchSpheres.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr long kActAlign = 64;  // floats: every activation region starts on a 256-byte boundary
constexpr long align_up(long v) { return (v + kActAlign - 1) / kActAlign * kActAlign; }
constexpr long kScratchSpheres = 1664;
constexpr long kScrZ1 = 0, kScrIn2 = kScrZ1 + align_up(kScratchSpheres * kNN * kH),
               kScrZ2 = kScrIn2 + align_up(kScratchSpheres * kIn2), kScrZ3 = kScrZ2 + align_up(kScratchSpheres * kH),
               kScrEnd = kScrZ3 + align_up(kScratchSpheres * kH);
static_assert(kScrEnd * (long)sizeof(float) <= (long)kScratchBytes, "activations exceed the scratch array");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};  // the array's device address on each device (nullptr: not looked up)

// with `dev` current: the device address of `scratch` on dev, looked up once per device
int scratch_on(int dev, float** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    g_gnn_error = "scratch: device index out of range";
    return IBEM3D_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      g_gnn_error = std::string("scratch: hipGetSymbolAddress: ") + hipGetErrorString(e);
      return IBEM3D_EHIP;
    }
    g_scratch[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch[dev];
  return 0;
}

void weight_grad(ibem3d_gnn* G, const float* X, int ldx, const float* dY, long M, int K, int N, float* gw, float* gb) {
  const long chunks = (M + kChunk - 1) / kChunk;
  const long len = (long)(K + 1) * N;
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(grid1(len), (unsigned)chunks), dim3(256), 0, G->st, X, ldx, dY, M, K, N,
                     G->part);
  hipLaunchKernelGGL(k_sum_chunks, dim3(grid1(len)), dim3(256), 0, G->st, G->part, chunks, len, gw, gb, K * N);
}

// loss and gradient (into G->g) of the minibatch already in b_edge / b_node / b_tgt; returns the
// number of loss partial sums written to G->lpart
int loss_grad(ibem3d_gnn* G, long B) {
  const long E = B * kNN;
  forward(G, G->b_edge, G->b_node, B, false);
  const long len = B * kOut;
  const int nb = (int)std::min<long>(1024, (len + 255) / 256);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, G->st, G->y, G->b_tgt, len, G->dy, G->lpart);
  float* P = G->p;
  float* g = G->g;
  weight_grad(G, G->z3, kH, G->dy, B, kH, kOut, g + kOffW4, g + kOffB4);
  hipLaunchKernelGGL(k_dense_bwd_x, dim3((unsigned)B), dim3(256), 0, G->st, G->dy, B, kOut, P + kOffW4, kH, G->z3,
                     G->dz3);
  weight_grad(G, G->z2, kH, G->dz3, B, kH, kH, g + kOffW3, g + kOffB3);
  hipLaunchKernelGGL(k_dense_bwd_x, dim3((unsigned)B), dim3(256), 0, G->st, G->dz3, B, kH, P + kOffW3, kH, G->z2,
                     G->dz2);
  weight_grad(G, G->in2, kIn2, G->dz2, B, kIn2, kH, g + kOffW2, g + kOffB2);
  hipLaunchKernelGGL(k_dense_bwd_x, dim3((unsigned)B), dim3(256), 0, G->st, G->dz2, B, kH, P + kOffW2, kIn2,
                     nullptr, G->din2);
  hipLaunchKernelGGL(k_aggregate_bwd, dim3((unsigned)B), dim3(128), 0, G->st, G->din2, G->b_edge, G->z1, B, G->dz1);
  weight_grad(G, G->b_edge, kFeS, G->dz1, E, kFe, kH, g + kOffW1, g + kOffB1);
  return nb;
}

// deterministic parameter initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from splitmix64
std::vector<float> init_params(unsigned long long seed) {
  std::vector<float> p(kParams, 0.0f);
  unsigned long long s = seed;
  auto next = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto fill = [&](long off, long count, int fan_in) {
    const double a = 1.0 / std::sqrt((double)fan_in);
    for (long u = 0; u < count; ++u) p[off + u] = (float)(a * (2.0 * ((next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
  };
  fill(kOffW1, kFe * kH, kFe);
  fill(kOffW2, kIn2 * kH, kIn2);
  fill(kOffW3, kH * kH, kH);
  fill(kOffW4, kH * kOut, kH);
  return p;
}

}  // namespace

extern "C" {

const char* ibem3d_gnn_last_error(void) { return g_gnn_error.c_str(); }
long ibem3d_gnn_param_count(void) { return kParams; }

int ibem3d_gnn_graph(const double* centres, const double* radii, int num, double E1_over_E0, double nu1,
                     int device, int* nbr, float* edge, float* node) {
  if (num <= 0 || !centres || !radii || !nbr || !edge || !node || !(E1_over_E0 > 0)) {
    g_gnn_error = "gnn_graph: invalid arguments";
    return IBEM3D_EINVAL;
  }
  GCHK(hipSetDevice(device));
  double *dc = nullptr, *dr = nullptr;
  int* dn = nullptr;
  float *de = nullptr, *df = nullptr;
  int rc = 0;
  auto cleanup = [&]() {
    hipFree(dc);
    hipFree(dr);
    hipFree(dn);
    hipFree(de);
    hipFree(df);
  };
  if (hipMalloc((void**)&dc, sizeof(double) * 3 * num) != hipSuccess ||
      hipMalloc((void**)&dr, sizeof(double) * num) != hipSuccess ||
      hipMalloc((void**)&dn, sizeof(int) * (long)num * kNN) != hipSuccess ||
      hipMalloc((void**)&de, sizeof(float) * (long)num * kNN * kFeS) != hipSuccess ||
      hipMalloc((void**)&df, sizeof(float) * (long)num * kFnS) != hipSuccess ||
      hipMemcpy(dc, centres, sizeof(double) * 3 * num, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dr, radii, sizeof(double) * num, hipMemcpyHostToDevice) != hipSuccess) {
    g_gnn_error = "gnn_graph: device buffers";
    cleanup();
    return IBEM3D_EHIP;
  }
  hipLaunchKernelGGL(k_knn, dim3(grid1(num)), dim3(256), 0, 0, dc, num, dn);
  hipLaunchKernelGGL(k_features, dim3(grid1(num)), dim3(256), 0, 0, dc, dr, num, dn,
                     (float)std::log(E1_over_E0), (float)nu1, de, df);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(nbr, dn, sizeof(int) * (long)num * kNN, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(edge, de, sizeof(float) * (long)num * kNN * kFeS, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(node, df, sizeof(float) * (long)num * kFnS, hipMemcpyDeviceToHost) != hipSuccess) {
    g_gnn_error = "gnn_graph: kernels or copies failed";
    rc = IBEM3D_EHIP;
  }
  cleanup();
  return rc;
}

ibem3d_gnn* ibem3d_gnn_create(int device, unsigned long long seed) {
  if (hipSetDevice(device) != hipSuccess) {
    g_gnn_error = "gnn_create: hipSetDevice";
    return nullptr;
  }
  ibem3d_gnn* G = new ibem3d_gnn();
  G->dev = device;
  const std::vector<float> p0 = init_params(seed);
  if (hipStreamCreateWithFlags(&G->st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&G->p, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc((void**)&G->g, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc((void**)&G->am, sizeof(float) * kParams) != hipSuccess ||
      hipMalloc((void**)&G->av, sizeof(float) * kParams) != hipSuccess ||
      hipMemcpy(G->p, p0.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess) {
    g_gnn_error = "gnn_create: device buffers";
    ibem3d_gnn_destroy(G);
    return nullptr;
  }
  std::vector<float> z(kParams, 0.0f);
  if (hipMemcpy(G->am, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(G->av, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) != hipSuccess) {
    g_gnn_error = "gnn_create: device buffers";
    ibem3d_gnn_destroy(G);
    return nullptr;
  }
  float* scr = nullptr;  // the no-allocation prediction's array, looked up here with the device current
  if (scratch_on(device, &scr)) {
    ibem3d_gnn_destroy(G);
    return nullptr;
  }
  return G;
}

void ibem3d_gnn_destroy(ibem3d_gnn* G) {
  if (!G) return;
  hipSetDevice(G->dev);
  if (G->st) hipStreamSynchronize(G->st);
  void* bufs[] = {G->p,    G->g,     G->am,    G->av,   G->d_edge, G->d_node, G->d_tgt, G->b_edge, G->b_node,
                  G->b_tgt, G->z1,   G->in2,   G->z2,   G->z3,     G->y,      G->dy,    G->dz3,    G->dz2,
                  G->din2, G->dz1,   G->part,  G->lpart, G->b_idx};
  for (void* q : bufs) hipFree(q);
  if (G->st) hipStreamDestroy(G->st);
  delete G;
}

int ibem3d_gnn_get_params(ibem3d_gnn* G, float* p) {
  if (!G || !p) return IBEM3D_EINVAL;
  GCHK(hipSetDevice(G->dev));
  GCHK(hipMemcpyAsync(p, G->p, sizeof(float) * kParams, hipMemcpyDeviceToHost, G->st));
  GCHK(hipStreamSynchronize(G->st));
  return 0;
}

int ibem3d_gnn_set_params(ibem3d_gnn* G, const float* p) {
  if (!G || !p) return IBEM3D_EINVAL;
  GCHK(hipSetDevice(G->dev));
  GCHK(hipMemcpyAsync(G->p, p, sizeof(float) * kParams, hipMemcpyHostToDevice, G->st));
  GCHK(hipStreamSynchronize(G->st));
  G->adam_t = 0;
  std::vector<float> z(kParams, 0.0f);
  GCHK(hipMemcpy(G->am, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(G->av, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  return 0;
}

int ibem3d_gnn_set_data(ibem3d_gnn* G, const float* edge, const float* node, const float* target, long S) {
  if (!G || !edge || !node || !target || S <= 0) {
    g_gnn_error = "gnn_set_data: invalid arguments";
    return IBEM3D_EINVAL;
  }
  GCHK(hipSetDevice(G->dev));
  hipFree(G->d_edge);
  hipFree(G->d_node);
  hipFree(G->d_tgt);
  G->d_edge = G->d_node = G->d_tgt = nullptr;
  GCHK(hipMalloc((void**)&G->d_edge, sizeof(float) * S * kNN * kFeS));
  GCHK(hipMalloc((void**)&G->d_node, sizeof(float) * S * kFnS));
  GCHK(hipMalloc((void**)&G->d_tgt, sizeof(float) * S * kOut));
  GCHK(hipMemcpy(G->d_edge, edge, sizeof(float) * S * kNN * kFeS, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(G->d_node, node, sizeof(float) * S * kFnS, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(G->d_tgt, target, sizeof(float) * S * kOut, hipMemcpyHostToDevice));
  G->S = S;
  return 0;
}

// loss and gradient over the dataset rows idx[0..B) (no parameter update)
int ibem3d_gnn_loss_grad(ibem3d_gnn* G, const int* idx, long B, float* loss, float* grad) {
  if (!G || !idx || B <= 0 || G->S <= 0) {
    g_gnn_error = "gnn_loss_grad: invalid arguments or no data";
    return IBEM3D_EINVAL;
  }
  for (long r = 0; r < B; ++r)
    if (idx[r] < 0 || idx[r] >= G->S) {
      g_gnn_error = "gnn_loss_grad: index out of range";
      return IBEM3D_EINVAL;
    }
  GCHK(hipSetDevice(G->dev));
  if (int rc = ensure_batch(G, B)) return rc;
  GCHK(hipMemcpyAsync(G->b_idx, idx, sizeof(int) * B, hipMemcpyHostToDevice, G->st));
  hipLaunchKernelGGL(k_gather, dim3(grid1(B * kNN * kFeS)), dim3(256), 0, G->st, G->b_idx, B, G->d_edge,
                     (long)kNN * kFeS, G->b_edge);
  hipLaunchKernelGGL(k_gather, dim3(grid1(B * kFnS)), dim3(256), 0, G->st, G->b_idx, B, G->d_node, (long)kFnS,
                     G->b_node);
  hipLaunchKernelGGL(k_gather, dim3(grid1(B * kOut)), dim3(256), 0, G->st, G->b_idx, B, G->d_tgt, (long)kOut,
                     G->b_tgt);
  const int nb = loss_grad(G, B);
  std::vector<float> lp(nb);
  GCHK(hipMemcpyAsync(lp.data(), G->lpart, sizeof(float) * nb, hipMemcpyDeviceToHost, G->st));
  if (grad) GCHK(hipMemcpyAsync(grad, G->g, sizeof(float) * kParams, hipMemcpyDeviceToHost, G->st));
  GCHK(hipStreamSynchronize(G->st));
  double acc = 0.0;
  for (float v : lp) acc += v;
  if (loss) *loss = (float)(acc / ((double)B * kOut));
  return 0;
}

// Adam training over the dataset: epochs passes over a permutation of the rows (perms: epochs x S
// row indices, the caller's shuffles), minibatches of B rows; loss_hist[e] = mean minibatch loss of
// epoch e (may be NULL)
int ibem3d_gnn_train(ibem3d_gnn* G, const int* perms, int epochs, long B, float lr, float* loss_hist) {
  if (!G || !perms || epochs <= 0 || B <= 0 || G->S <= 0 || !(lr > 0)) {
    g_gnn_error = "gnn_train: invalid arguments or no data";
    return IBEM3D_EINVAL;
  }
  GCHK(hipSetDevice(G->dev));
  if (int rc = ensure_batch(G, B)) return rc;
  int* d_perm = nullptr;
  GCHK(hipMalloc((void**)&d_perm, sizeof(int) * G->S));
  std::vector<float> lp(1024);
  int rc = 0;
  for (int ep = 0; ep < epochs && rc == 0; ++ep) {
    const int* perm = perms + (long)ep * G->S;
    for (long r = 0; r < G->S; ++r)
      if (perm[r] < 0 || perm[r] >= G->S) {
        g_gnn_error = "gnn_train: permutation index out of range";
        rc = IBEM3D_EINVAL;
        break;
      }
    if (rc) break;
    if (hipMemcpyAsync(d_perm, perm, sizeof(int) * G->S, hipMemcpyHostToDevice, G->st) != hipSuccess) {
      rc = IBEM3D_EHIP;
      break;
    }
    double ep_loss = 0.0;
    long steps = 0;
    for (long s0 = 0; s0 + B <= G->S; s0 += B) {
      const int* idx = d_perm + s0;
      hipLaunchKernelGGL(k_gather, dim3(grid1(B * kNN * kFeS)), dim3(256), 0, G->st, idx, B, G->d_edge,
                         (long)kNN * kFeS, G->b_edge);
      hipLaunchKernelGGL(k_gather, dim3(grid1(B * kFnS)), dim3(256), 0, G->st, idx, B, G->d_node, (long)kFnS,
                         G->b_node);
      hipLaunchKernelGGL(k_gather, dim3(grid1(B * kOut)), dim3(256), 0, G->st, idx, B, G->d_tgt, (long)kOut,
                         G->b_tgt);
      const int nb = loss_grad(G, B);
      ++G->adam_t;
      const float c1 = 1.0f - std::pow(0.9f, (float)G->adam_t), c2 = 1.0f - std::pow(0.999f, (float)G->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(grid1(kParams)), dim3(256), 0, G->st, G->p, G->g, G->am, G->av, kParams, lr,
                         c1, c2);
      if (loss_hist) {
        if (hipMemcpyAsync(lp.data(), G->lpart, sizeof(float) * nb, hipMemcpyDeviceToHost, G->st) != hipSuccess ||
            hipStreamSynchronize(G->st) != hipSuccess) {
          rc = IBEM3D_EHIP;
          break;
        }
        double acc = 0.0;
        for (int u = 0; u < nb; ++u) acc += lp[u];
        ep_loss += acc / ((double)B * kOut);
      }
      ++steps;
    }
    if (loss_hist) loss_hist[ep] = steps ? (float)(ep_loss / steps) : 0.0f;
  }
  if (hipStreamSynchronize(G->st) != hipSuccess || hipGetLastError() != hipSuccess) {
    if (!rc) g_gnn_error = "gnn_train: kernel failure";
    rc = rc ? rc : IBEM3D_EHIP;
  }
  hipFree(d_perm);
  return rc;
}

// prediction y (S x 360) for S spheres' edges (S x 16 x 8) and nodes (S x 4); bf16 = 1: the dense
// layers in BF16 on MFMA (FP32 accumulation), else FP32
int ibem3d_gnn_predict(ibem3d_gnn* G, const float* edge, const float* node, long S, int bf16, float* y) {
  if (!G || !edge || !node || !y || S <= 0) {
    g_gnn_error = "gnn_predict: invalid arguments";
    return IBEM3D_EINVAL;
  }
  GCHK(hipSetDevice(G->dev));
  constexpr long kPredChunk = 4096;  // spheres per pass (the edge layer's grid stays in range)
  if (int rc = ensure_batch(G, std::min(S, kPredChunk))) return rc;
  for (long s0 = 0; s0 < S; s0 += kPredChunk) {
    const long c = std::min(kPredChunk, S - s0);
    GCHK(hipMemcpyAsync(G->b_edge, edge + s0 * kNN * kFeS, sizeof(float) * c * kNN * kFeS, hipMemcpyHostToDevice,
                        G->st));
    GCHK(hipMemcpyAsync(G->b_node, node + s0 * kFnS, sizeof(float) * c * kFnS, hipMemcpyHostToDevice, G->st));
    forward(G, G->b_edge, G->b_node, c, bf16 != 0);
    GCHK(hipGetLastError());
    GCHK(hipMemcpyAsync(y + s0 * kOut, G->y, sizeof(float) * c * kOut, hipMemcpyDeviceToHost, G->st));
  }
  GCHK(hipStreamSynchronize(G->st));
  return 0;
}

// ibem3d_gnn_predict on device buffers with no allocation: the activations in the device's scratch
// array, the launches enqueued on `stream` (NULL: g's stream) and not waited for
int ibem3d_gnn_predict_noalloc(ibem3d_gnn* G, const float* d_edge, const float* d_node, long S, int bf16,
                               float* d_y, void* stream) {
  if (!G || !d_edge || !d_node || !d_y || S <= 0) {
    g_gnn_error = "gnn_predict_noalloc: invalid arguments";
    return IBEM3D_EINVAL;
  }
  GCHK(hipSetDevice(G->dev));
  float* scr = nullptr;
  if (int rc = scratch_on(G->dev, &scr)) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : G->st;
  for (long s0 = 0; s0 < S; s0 += kScratchSpheres) {
    const long c = std::min(kScratchSpheres, S - s0);
    forward_into(st, G->p, d_edge + s0 * kNN * kFeS, d_node + s0 * kFnS, c, bf16 != 0, scr + kScrZ1,
                 scr + kScrIn2, scr + kScrZ2, scr + kScrZ3, d_y + s0 * kOut);
  }
  GCHK(hipGetLastError());
  return 0;
}

}  // extern "C"
