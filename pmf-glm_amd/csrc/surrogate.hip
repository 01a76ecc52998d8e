// The learned face flux: data dump from the plain solver, training kernels (bfloat16
// operands on MFMA with float32 accumulation, float32 master weights, Adam), batch
// evaluation, and the host entry points (declared in include/pmfmhd.h).
//
// Network: 16 -> 64 -> 64 -> 64 -> 9, tanh-form GELU, y = W4^T g(W3^T g(W2^T g(W1^T x +
// b1) + b2) + b3) + b4. Weights are stored W_l[in][out] row-major; the flat parameter
// vector is [W1 b1 W2 b2 W3 b3 W4 b4] (9993 floats). Encoding of x and y: surrogate.hpp.
#include <cmath>
#include <cstring>
#include <random>

#include "cellpredict.hpp"
#include "state.hpp"
#include "surrogate.hpp"

namespace pmf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NPARAM = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 * 9 + 9;  // 9993
constexpr int LAYER_IN[4] = {16, 64, 64, 64};
constexpr int LAYER_OUT[4] = {64, 64, 64, 9};

__host__ __device__ constexpr int w_off(int l) {
  return l == 0 ? 0 : l == 1 ? 16 * 64 + 64 : l == 2 ? 16 * 64 + 64 + 64 * 64 + 64 : 16 * 64 + 64 + 2 * (64 * 64 + 64);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + LAYER_IN[l] * LAYER_OUT[l]; }

// ---------------------------------------------------------------------------
// C[M x N] (+)= op(A)[M x K] op(B)[K x N], row-major storage; op(A) = A (M x K, lda) or
// A^T (A stored K x M); likewise B. One 64x64 tile per workgroup, one 32x32 MFMA tile per
// wave. BF16: operands rounded to bfloat16, v_mfma_f32_32x32x8_bf16; otherwise float32
// operands on v_mfma_f32_32x32x2_f32. Lane layout (32x32 MFMA, wave64): A[i][k] from lane
// i + 32*(k/kb), B[k][j] from lane j + 32*(k/kb) (kb = 4 or 1 per half), D[i][j] in lane
// j + 32*((i/4)%2), register 4*(i/8) + i%4.
template <bool TA, bool TB, bool BF16>
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, const float* __restrict__ B,
                                              float* __restrict__ C, int M, int N, int K, int lda, int ldb, int ldc,
                                              int accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i0 = blockIdx.x * 64 + (wave & 1) * 32, j0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int li = lane & 31, lh = lane >> 5;
  auto getA = [&](int i, int k) -> float {
    if (i >= M || k >= K) return 0.f;
    return TA ? A[(int64_t)k * lda + i] : A[(int64_t)i * lda + k];
  };
  auto getB = [&](int k, int j) -> float {
    if (j >= N || k >= K) return 0.f;
    return TB ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j];
  };
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  if constexpr (BF16) {
    for (int k0 = 0; k0 < K; k0 += 8) {
      bf16x4 a, b;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = (short)bf16_bits(getA(i0 + li, k0 + 4 * lh + q));
        b[q] = (short)bf16_bits(getB(k0 + 4 * lh + q, j0 + li));
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
  } else {
    for (int k0 = 0; k0 < K; k0 += 2) {
      float a = getA(i0 + li, k0 + lh), b = getB(k0 + lh, j0 + li);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = i0 + 8 * (r >> 2) + 4 * lh + (r & 3), j = j0 + li;
    if (i < M && j < N) {
      float* c = &C[(int64_t)i * ldc + j];
      *c = accumulate ? *c + acc[r] : acc[r];
    }
  }
}

static void gemm(bool ta, bool tb, bool bf16, const float* A, const float* B, float* C, int M, int N, int K, int lda,
                 int ldb, int ldc, bool acc, hipStream_t st) {
  dim3 grd((M + 63) / 64, (N + 63) / 64), blk(256);
#define PMF_G(a, b, c) hipLaunchKernelGGL((k_gemm<a, b, c>), grd, blk, 0, st, A, B, C, M, N, K, lda, ldb, ldc, acc ? 1 : 0)
  if (bf16) {
    if (!ta && !tb) PMF_G(false, false, true);
    else if (ta && !tb) PMF_G(true, false, true);
    else if (!ta && tb) PMF_G(false, true, true);
    else PMF_G(true, true, true);
  } else {
    if (!ta && !tb) PMF_G(false, false, false);
    else if (ta && !tb) PMF_G(true, false, false);
    else if (!ta && tb) PMF_G(false, true, false);
    else PMF_G(true, true, false);
  }
#undef PMF_G
}

// Z[r][c] += b[c]; A = GELU(Z) (or A = Z for the output layer).
__global__ __launch_bounds__(256) void k_bias_act(float* __restrict__ Z, const float* __restrict__ b,
                                                  float* __restrict__ A, int64_t rows, int cols, int act) {
  int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= rows * cols) return;
  float z = Z[e] + b[e % cols];
  Z[e] = z;
  if (A) A[e] = act ? gelu(z) : z;
}

// dY = 2 (Y - T) / (rows * 7) on the learned slots (0-4, 6, 7); per-block partial loss.
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ Y, const float* __restrict__ Tg,
                                                   float* __restrict__ dY, int64_t rows, double* __restrict__ part) {
  __shared__ double sh[4];
  int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (e < rows * 9) {
    int k = (int)(e % 9);
    bool learned = k != 5 && k != 8;
    float d = learned ? Y[e] - Tg[e] : 0.f;
    dY[e] = 2.f * d / float(rows * 7);
    l = (double)d * d;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void k_sum_flat(const double* __restrict__ part, int n, double scale,
                                                  double* __restrict__ out) {
  __shared__ double sh[4];
  double l = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) l += part[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) *out = (sh[0] + sh[1] + sh[2] + sh[3]) * scale;
}

// dZ = dA * GELU'(Z)
__global__ __launch_bounds__(256) void k_gelu_bwd(const float* __restrict__ dA, const float* __restrict__ Z,
                                                  float* __restrict__ dZ, int64_t n) {
  int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) dZ[e] = dA[e] * gelu_grad(Z[e]);
}

// db[c] = sum_r dZ[r][c] (one block per column, float64 accumulation)
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, int64_t rows, int cols,
                                                float* __restrict__ db) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int64_t r = threadIdx.x; r < rows; r += 256) s += dZ[r * cols + c];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) db[c] = (float)(sh[0] + sh[1] + sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  float gi = g[e];
  float mi = b1 * m[e] + (1.f - b1) * gi;
  float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  w[e] -= lr * (mi * c1) / (__builtin_sqrtf(vi * c2) + eps);
}

// Mini-batch gather with hashed indices (sampling with replacement).
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ X, const float* __restrict__ Tg, int64_t n,
                                                float* __restrict__ Xb, float* __restrict__ Tb, int batch,
                                                unsigned long long key) {
  int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= batch) return;
  unsigned long long z = key + (unsigned long long)r * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  int64_t i = (int64_t)(z % (unsigned long long)n);
#pragma unroll
  for (int k = 0; k < 16; ++k) Xb[(int64_t)r * 16 + k] = X[i * 16 + k];
#pragma unroll
  for (int k = 0; k < 9; ++k) Tb[(int64_t)r * 9 + k] = Tg[i * 9 + k];
}

// ---------------------------------------------------------------------------
// Face dump: every lower face of every cell (x, y, z) of the current state, predicted
// exactly as the per-cell step does; gated faces are kept with probability keep and
// appended as (x[16], y[9], state[18]) where state = normal-frame L[8], R[8], fbn, fpsi.
__global__ __launch_bounds__(256) void k_dump_faces(Fields<const float> U, Grid G, StepParams<float> P, float thr,
                                                    float keep, unsigned long long key, float* __restrict__ X,
                                                    float* __restrict__ Y, float* __restrict__ S,
                                                    unsigned long long* __restrict__ count, int64_t cap) {
  int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.nc) return;
  const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  float mh[9], dL[3][9], dR[3][9], mhn[9], dLn[3][9], dRn[3][9];
  predict_cell<RECON_PLM_MINMOD>(U, G, i, j, k, P, mh, dL, dR);
  for (int d = 0; d < 3; ++d) {
    unsigned long long z = key ^ ((unsigned long long)c * 3 + d) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if ((float)(z >> 40) * (1.0f / 16777216.0f) >= keep) continue;
    predict_cell<RECON_PLM_MINMOD>(U, G, i - (d == 0), j - (d == 1), k - (d == 2), P, mhn, dLn, dRn);
    float Lq[9], Rq[9];
    for (int v = 0; v < 9; ++v) {
      Lq[v] = mhn[v] + dRn[d][v];
      Rq[v] = mh[v] + dL[d][v];
    }
    const int N = d, A = (d + 1) % 3, B = (d + 2) % 3;
    float L[9] = {Lq[0], Lq[1 + N], Lq[1 + A], Lq[1 + B], Lq[4], Lq[5 + N], Lq[5 + A], Lq[5 + B], Lq[8]};
    float R[9] = {Rq[0], Rq[1 + N], Rq[1 + A], Rq[1 + B], Rq[4], Rq[5 + N], Rq[5 + A], Rq[5 + B], Rq[8]};
    float fbn, fpsi;
    bool uself = surr_prepare(L, R, P, fbn, fpsi);
    float x[16], s[9];
    bool smooth = surr_encode(L, R, P.gamma, thr, x, s);
    if (!smooth || uself) continue;
    float fh[9], fc[9];
    hlld_n(L, R, L[5], fbn, fpsi, fh, P);
    surr_central(L, R, P.igm1, fc);
    unsigned long long slot = atomicAdd(count, 1ull);
    if ((int64_t)slot >= cap) continue;
    for (int q = 0; q < 16; ++q) X[slot * 16 + q] = x[q];
    for (int q = 0; q < 9; ++q) Y[slot * 9 + q] = (q == 5 || q == 8) ? 0.f : (fh[q] - fc[q]) / s[q];
    for (int q = 0; q < 8; ++q) {
      S[slot * 18 + q] = L[q];
      S[slot * 18 + 8 + q] = R[q];
    }
    S[slot * 18 + 16] = fbn;
    S[slot * 18 + 17] = fpsi;
  }
}

// Face fluxes of dumped states: mode 0 = HLLD, 1 = central + thr s y (surrogate), 2 =
// central only (the method without the learned correction).
__global__ __launch_bounds__(256) void k_face_flux(const float* __restrict__ S, const float* __restrict__ Yp, int64_t n,
                                                   StepParams<float> P, float thr, int mode, float* __restrict__ F) {
  int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= n) return;
  float L[9], R[9];
  for (int q = 0; q < 8; ++q) {
    L[q] = S[f * 18 + q];
    R[q] = S[f * 18 + 8 + q];
  }
  L[8] = R[8] = 0.f;
  const float fbn = S[f * 18 + 16], fpsi = S[f * 18 + 17];
  float out[9];
  if (mode == 0) {
    hlld_n(L, R, L[5], fbn, fpsi, out, P);
  } else {
    float x[16], s[9], fc[9];
    surr_encode(L, R, P.gamma, thr, x, s);
    surr_central(L, R, P.igm1, fc);
    for (int q = 0; q < 9; ++q) out[q] = fc[q] + (mode == 1 ? s[q] * Yp[f * 9 + q] : 0.f);
    out[5] = fbn;
    out[8] = fpsi;
  }
  for (int q = 0; q < 9; ++q) F[f * 9 + q] = out[q];
}

// The network's LDS image for the surrogate-assisted step (layout in surrogate.hpp):
// transposed bfloat16 weights with padded rows, then the float32 biases.
__global__ __launch_bounds__(256) void k_surr_image(const float* __restrict__ w, unsigned char* __restrict__ img) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  unsigned short* W = reinterpret_cast<unsigned short*>(img);
  float* B = reinterpret_cast<float*>(img + SIMG_NW * 2);
  if (e < SIMG_NW) {
    float val = 0.f;
    if (e < SIMG_W2) {
      int r = e / SIMG_S1, c = e % SIMG_S1;
      if (c < 16) val = w[w_off(0) + c * 64 + r];
    } else if (e < SIMG_W4) {
      int l = e < SIMG_W3 ? 1 : 2, o = e - (l == 1 ? SIMG_W2 : SIMG_W3);
      int r = o / SIMG_S2, c = o % SIMG_S2;
      if (c < 64) val = w[w_off(l) + c * 64 + r];
    } else {
      int o = e - SIMG_W4, r = o / SIMG_S2, c = o % SIMG_S2;
      if (r < 9 && c < 64) val = w[w_off(3) + c * 9 + r];
    }
    W[e] = bf16_bits(val);
  }
  if (e < SIMG_NB) {
    float val = 0.f;
    if (e < 192) val = w[b_off(e / 64) + e % 64];
    else if (e - 192 < 9) val = w[b_off(3) + e - 192];
    B[e] = val;
  }
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

}  // namespace pmf

// ---------------------------------------------------------------------------
// Host side.
using namespace pmf;

struct pmf_surrogate {
  pmf_state* s = nullptr;  // parameters and stream (not owned)
  float* w = nullptr;      // NPARAM master weights
  float* g = nullptr;      // gradient
  float* m = nullptr;      // Adam moments
  float* v = nullptr;
  int64_t adam_t = 0;
  // dataset
  float *X = nullptr, *Y = nullptr, *S = nullptr;
  int64_t cap = 0, n = 0;
  unsigned long long* dcount = nullptr;
  // activations of one batch
  int64_t act_rows = 0;
  float *Z[4] = {nullptr, nullptr, nullptr, nullptr}, *A[4] = {nullptr, nullptr, nullptr, nullptr};
  float *dZ = nullptr, *dA = nullptr, *Xb = nullptr, *Tb = nullptr;
  double *part = nullptr, *dloss = nullptr, *hloss = nullptr;
  unsigned char* img = nullptr;         // LDS image for the surrogate-assisted step
  unsigned long long* nsmooth = nullptr;  // faces solved by the network in the last step
};

namespace {

void free_acts(pmf_surrogate* n) {
  for (int l = 0; l < 4; ++l) {
    (void)hipFree(n->Z[l]);
    (void)hipFree(n->A[l]);
    n->Z[l] = n->A[l] = nullptr;
  }
  (void)hipFree(n->dZ);
  (void)hipFree(n->dA);
  (void)hipFree(n->Xb);
  (void)hipFree(n->Tb);
  (void)hipFree(n->part);
  n->dZ = n->dA = n->Xb = n->Tb = nullptr;
  n->part = nullptr;
  n->act_rows = 0;
}

void ensure_acts(pmf_surrogate* n, int64_t rows) {
  if (rows <= n->act_rows) return;
  free_acts(n);
  for (int l = 0; l < 4; ++l) {
    hip_check(hipMalloc(&n->Z[l], sizeof(float) * rows * 64), "hipMalloc");
    hip_check(hipMalloc(&n->A[l], sizeof(float) * rows * 64), "hipMalloc");
  }
  hip_check(hipMalloc(&n->dZ, sizeof(float) * rows * 64), "hipMalloc");
  hip_check(hipMalloc(&n->dA, sizeof(float) * rows * 64), "hipMalloc");
  hip_check(hipMalloc(&n->Xb, sizeof(float) * rows * 16), "hipMalloc");
  hip_check(hipMalloc(&n->Tb, sizeof(float) * rows * 9), "hipMalloc");
  hip_check(hipMalloc(&n->part, sizeof(double) * ((rows * 9 + 255) / 256 + 1)), "hipMalloc");
  n->act_rows = rows;
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// Forward pass of rows samples X (row-major rows x 16): Z[l], A[l] filled, output in Z[3].
void forward(pmf_surrogate* n, const float* X, int64_t rows, bool bf16) {
  hipStream_t st = n->s->st;
  const float* in = X;
  for (int l = 0; l < 4; ++l) {
    const int K = LAYER_IN[l], N = LAYER_OUT[l];
    gemm(false, false, bf16, in, n->w + w_off(l), n->Z[l], (int)rows, N, K, K, N, N, false, st);
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(rows * N)), dim3(256), 0, st, n->Z[l], n->w + b_off(l),
                       l < 3 ? n->A[l] : (float*)nullptr, rows, N, l < 3 ? 1 : 0);
    in = n->A[l];
  }
}

// Gradient of the mean-squared loss over rows samples into n->g; loss into n->dloss.
void backward(pmf_surrogate* n, const float* X, const float* Tg, int64_t rows, bool bf16) {
  hipStream_t st = n->s->st;
  const unsigned nb = blocks(rows * 9);
  hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, st, n->Z[3], Tg, n->dZ, rows, n->part);
  hipLaunchKernelGGL(k_sum_flat, dim3(1), dim3(256), 0, st, n->part, (int)nb, 1.0 / double(rows * 7), n->dloss);
  // layer 4: dZ holds dY (rows x 9)
  for (int l = 3; l >= 0; --l) {
    const int K = LAYER_IN[l], N = LAYER_OUT[l];
    const float* Ain = l == 0 ? X : n->A[l - 1];
    // dW_l[K x N] = Ain^T dZ
    gemm(true, false, bf16, Ain, n->dZ, n->g + w_off(l), K, N, (int)rows, K, N, N, false, st);
    hipLaunchKernelGGL(k_colsum, dim3(N), dim3(256), 0, st, n->dZ, rows, N, n->g + b_off(l));
    if (l > 0) {
      // dA_{l-1}[rows x K] = dZ W_l^T ; dZ_{l-1} = dA * GELU'(Z_{l-1})
      gemm(false, true, bf16, n->dZ, n->w + w_off(l), n->dA, (int)rows, K, N, N, N, K, false, st);
      hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks(rows * K)), dim3(256), 0, st, n->dA, n->Z[l - 1], n->dZ, rows * K);
    }
  }
}

}  // namespace

extern "C" {

int pmf_surrogate_create(pmf_state* s, uint64_t seed, pmf_surrogate** out) {
  return guard([&] {
    check_state(s);
    if (!out) throw Error("null argument");
    if (s->p.precision != PMF_F32) throw Error("the flux surrogate runs on float32 states");
    auto* n = new pmf_surrogate();
    try {
      n->s = s;
      hip_check(hipMalloc(&n->w, sizeof(float) * NPARAM), "hipMalloc");
      hip_check(hipMalloc(&n->g, sizeof(float) * NPARAM), "hipMalloc");
      hip_check(hipMalloc(&n->m, sizeof(float) * NPARAM), "hipMalloc");
      hip_check(hipMalloc(&n->v, sizeof(float) * NPARAM), "hipMalloc");
      hip_check(hipMalloc(&n->dloss, sizeof(double)), "hipMalloc");
      hip_check(hipMalloc(&n->dcount, sizeof(unsigned long long)), "hipMalloc");
      hip_check(hipMalloc(&n->img, SIMG_BYTES), "hipMalloc");
      hip_check(hipMalloc(&n->nsmooth, sizeof(unsigned long long)), "hipMalloc");
      hip_check(hipHostMalloc((void**)&n->hloss, sizeof(double) * 2, hipHostMallocDefault), "hipHostMalloc");
      // He-style uniform init of the weights, zero biases
      std::vector<float> h(NPARAM, 0.f);
      std::mt19937_64 rng(seed);
      for (int l = 0; l < 4; ++l) {
        float lim = std::sqrt(6.0f / float(LAYER_IN[l] + LAYER_OUT[l])) * (l < 3 ? 1.0f : 0.1f);
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int e = 0; e < LAYER_IN[l] * LAYER_OUT[l]; ++e) h[w_off(l) + e] = u(rng);
      }
      hip_check(hipMemcpy(n->w, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice), "copy");
      hip_check(hipMemset(n->m, 0, sizeof(float) * NPARAM), "memset");
      hip_check(hipMemset(n->v, 0, sizeof(float) * NPARAM), "memset");
      // null-stream work: the state's non-blocking stream does not wait for it
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    } catch (...) {
      pmf_surrogate_destroy(n);
      throw;
    }
    *out = n;
  });
}

int pmf_surrogate_destroy(pmf_surrogate* n) {
  return guard([&] {
    if (!n) return;
    if (n->s) (void)hipStreamSynchronize(n->s->st);
    free_acts(n);
    (void)hipFree(n->w);
    (void)hipFree(n->g);
    (void)hipFree(n->m);
    (void)hipFree(n->v);
    (void)hipFree(n->X);
    (void)hipFree(n->Y);
    (void)hipFree(n->S);
    (void)hipFree(n->dloss);
    (void)hipFree(n->dcount);
    (void)hipFree(n->img);
    (void)hipFree(n->nsmooth);
    if (n->hloss) (void)hipHostFree(n->hloss);
    delete n;
  });
}

int pmf_surrogate_nparams(void) { return NPARAM; }

int pmf_surrogate_get_weights(pmf_surrogate* n, float* host, int64_t count) {
  return guard([&] {
    if (!n || !host || count != NPARAM) throw Error("weights array must hold pmf_surrogate_nparams() floats");
    hip_check(hipMemcpyAsync(host, n->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, n->s->st), "copy");
    sync(n->s);
  });
}

int pmf_surrogate_set_weights(pmf_surrogate* n, const float* host, int64_t count) {
  return guard([&] {
    if (!n || !host || count != NPARAM) throw Error("weights array must hold pmf_surrogate_nparams() floats");
    hip_check(hipMemcpyAsync(n->w, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, n->s->st), "copy");
    sync(n->s);
  });
}

int pmf_surrogate_reserve(pmf_surrogate* n, int64_t cap) {
  return guard([&] {
    if (!n || cap < 1) throw Error("bad arguments");
    (void)hipFree(n->X);
    (void)hipFree(n->Y);
    (void)hipFree(n->S);
    n->X = n->Y = n->S = nullptr;
    hip_check(hipMalloc(&n->X, sizeof(float) * 16 * cap), "hipMalloc(dataset)");
    hip_check(hipMalloc(&n->Y, sizeof(float) * 9 * cap), "hipMalloc(dataset)");
    hip_check(hipMalloc(&n->S, sizeof(float) * 18 * cap), "hipMalloc(dataset)");
    n->cap = cap;
    n->n = 0;
  });
}

int pmf_surrogate_dump(pmf_surrogate* n, pmf_state* s, double dt, double ch, double thr, double keep, uint64_t seed,
                       int64_t* total) {
  return guard([&] {
    if (!n || !n->X) throw Error("reserve the dataset first (pmf_surrogate_reserve)");
    check_fields(s);
    if (s->p.precision != PMF_F32) throw Error("the flux surrogate runs on float32 states");
    if (!(ch > 0)) throw Error("step! needs the GLM cleaning speed `ch` (GLM states)");
    unsigned long long c0 = (unsigned long long)n->n;
    hip_check(hipMemcpyAsync(n->dcount, &c0, sizeof(c0), hipMemcpyHostToDevice, s->st), "copy");
    StepParams<float> P = make_params<float>(s, dt, ch, 1.0);
    hipLaunchKernelGGL(k_dump_faces, dim3(blocks(s->nc)), dim3(256), 0, s->st, cfields_at<float>(s->U, s->nc), s->G,
                       P, float(thr), float(keep), (unsigned long long)seed, n->X, n->Y, n->S, n->dcount, n->cap);
    hip_check(hipGetLastError(), "dump launch");
    unsigned long long c1 = 0;
    hip_check(hipMemcpyAsync(&c1, n->dcount, sizeof(c1), hipMemcpyDeviceToHost, s->st), "copy");
    sync(s);
    n->n = std::min<int64_t>((int64_t)c1, n->cap);
    if (total) *total = n->n;
  });
}

int pmf_surrogate_dataset(pmf_surrogate* n, int64_t first, int64_t count, float* X, float* Y, float* S) {
  return guard([&] {
    if (!n || first < 0 || count < 0 || first + count > n->n) throw Error("dataset range out of bounds");
    hipStream_t st = n->s->st;
    if (X) hip_check(hipMemcpyAsync(X, n->X + first * 16, sizeof(float) * 16 * count, hipMemcpyDeviceToHost, st), "copy");
    if (Y) hip_check(hipMemcpyAsync(Y, n->Y + first * 9, sizeof(float) * 9 * count, hipMemcpyDeviceToHost, st), "copy");
    if (S) hip_check(hipMemcpyAsync(S, n->S + first * 18, sizeof(float) * 18 * count, hipMemcpyDeviceToHost, st), "copy");
    sync(n->s);
  });
}

int pmf_surrogate_upload(pmf_surrogate* n, const float* X, const float* Y, const float* S, int64_t count) {
  return guard([&] {
    if (!n || !X || !Y || count < 1) throw Error("bad arguments");
    if (count > n->cap) throw Error("dataset capacity exceeded");
    hipStream_t st = n->s->st;
    hip_check(hipMemcpyAsync(n->X, X, sizeof(float) * 16 * count, hipMemcpyHostToDevice, st), "copy");
    hip_check(hipMemcpyAsync(n->Y, Y, sizeof(float) * 9 * count, hipMemcpyHostToDevice, st), "copy");
    if (S) hip_check(hipMemcpyAsync(n->S, S, sizeof(float) * 18 * count, hipMemcpyHostToDevice, st), "copy");
    sync(n->s);
    n->n = count;
  });
}

int pmf_surrogate_grad(pmf_surrogate* n, int64_t first, int64_t count, int32_t bf16, float* grad, double* loss) {
  return guard([&] {
    if (!n || first < 0 || count < 1 || first + count > n->n) throw Error("dataset range out of bounds");
    ensure_acts(n, count);
    forward(n, n->X + first * 16, count, bf16 != 0);
    backward(n, n->X + first * 16, n->Y + first * 9, count, bf16 != 0);
    hip_check(hipGetLastError(), "surrogate launch");
    if (grad) hip_check(hipMemcpyAsync(grad, n->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, n->s->st), "copy");
    hip_check(hipMemcpyAsync(n->hloss, n->dloss, sizeof(double), hipMemcpyDeviceToHost, n->s->st), "copy");
    sync(n->s);
    if (loss) *loss = n->hloss[0];
  });
}

int pmf_surrogate_predict(pmf_surrogate* n, int64_t first, int64_t count, int32_t bf16, float* Yhost) {
  return guard([&] {
    if (!n || first < 0 || count < 1 || first + count > n->n) throw Error("dataset range out of bounds");
    ensure_acts(n, count);
    forward(n, n->X + first * 16, count, bf16 != 0);
    hip_check(hipGetLastError(), "surrogate launch");
    // Z[3] is rows x 9 with a leading dimension of 9
    hip_check(hipMemcpyAsync(Yhost, n->Z[3], sizeof(float) * 9 * count, hipMemcpyDeviceToHost, n->s->st), "copy");
    sync(n->s);
  });
}

int pmf_surrogate_train(pmf_surrogate* n, int32_t batch, int32_t steps, double lr, uint64_t seed, int32_t bf16,
                        double* losses) {
  return guard([&] {
    if (!n || n->n < 1) throw Error("the dataset is empty");
    if (batch < 1 || steps < 0) throw Error("bad arguments");
    ensure_acts(n, batch);
    hipStream_t st = n->s->st;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < steps; ++it) {
      unsigned long long key = seed * 0xD1B54A32D192ED03ull + (unsigned long long)(n->adam_t + 1) * 0x9E3779B97F4A7C15ull;
      hipLaunchKernelGGL(k_gather, dim3(blocks(batch)), dim3(256), 0, st, n->X, n->Y, n->n, n->Xb, n->Tb, batch, key);
      forward(n, n->Xb, batch, bf16 != 0);
      backward(n, n->Xb, n->Tb, batch, bf16 != 0);
      ++n->adam_t;
      float c1 = 1.f / (1.f - std::pow(b1, (float)n->adam_t));
      float c2 = 1.f / (1.f - std::pow(b2, (float)n->adam_t));
      hipLaunchKernelGGL(k_adam, dim3(blocks(NPARAM)), dim3(256), 0, st, n->w, n->g, n->m, n->v, NPARAM, float(lr), b1,
                         b2, eps, c1, c2);
      if (losses) {
        hip_check(hipMemcpyAsync(n->hloss, n->dloss, sizeof(double), hipMemcpyDeviceToHost, st), "copy");
        sync(n->s);
        losses[it] = n->hloss[0];
      }
    }
    hip_check(hipGetLastError(), "surrogate launch");
    sync(n->s);
  });
}

int pmf_surrogate_face_flux(pmf_surrogate* n, int64_t first, int64_t count, int32_t mode, double thr, float* Fhost) {
  return guard([&] {
    if (!n || first < 0 || count < 1 || first + count > n->n) throw Error("dataset range out of bounds");
    if (mode < 0 || mode > 2) throw Error("mode must be 0 (HLLD), 1 (surrogate) or 2 (central)");
    ensure_acts(n, count);
    hipStream_t st = n->s->st;
    if (mode == 1) forward(n, n->X + first * 16, count, true);
    float* F = n->dA;  // rows x 64 >= rows x 9
    StepParams<float> P = make_params<float>(n->s, 0.0, 1.0, 1.0);
    hipLaunchKernelGGL(k_face_flux, dim3(blocks(count)), dim3(256), 0, st, n->S + first * 18,
                       mode == 1 ? n->Z[3] : (const float*)nullptr, count, P, float(thr), mode, F);
    hip_check(hipGetLastError(), "face flux launch");
    hip_check(hipMemcpyAsync(Fhost, F, sizeof(float) * 9 * count, hipMemcpyDeviceToHost, st), "copy");
    sync(n->s);
  });
}

int pmf_step_surrogate(pmf_state* s, pmf_surrogate* n, double dt, double ch, double glm_cr, double thr,
                       int64_t* nsmooth) {
  return guard([&] {
    check_fields(s);
    if (!n) throw Error("null surrogate");
    if (s->p.precision != PMF_F32) throw Error("the flux surrogate runs on float32 states");
    if (s->p.solver != PMF_SOLVER_GLM) throw Error("the flux surrogate replaces the GLM-HLLD flux (GLM states only)");
    if (s->p.riemann != PMF_RIEMANN_HLLD) throw Error("the flux surrogate stands in for HLLD (riemann=:hlld)");
    if (!(ch > 0.0)) throw Error("step! needs the GLM cleaning speed `ch` (GLM states)");
    if (!surrogate_grid(s->G))
      throw Error("the surrogate step needs an all-periodic grid whose nx is a multiple of 16");
    const double decay = decay_of(s, dt, ch, glm_cr);
    hipLaunchKernelGGL(k_surr_image, dim3((SIMG_NW + 255) / 256), dim3(256), 0, s->st, n->w, n->img);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s->st, n->nsmooth);
    mark(s);
    bool ok = launch_step_fused_surr(cfields_at<float>(s->U, s->nc), fields_at<float>(s->S, s->nc), s->G,
                                     make_params<float>(s, dt, ch, decay), s->red, s->dval, n->img, float(thr),
                                     n->nsmooth, s->st);
    mark(s);
    if (!ok) throw Error("surrogate step unavailable for this configuration");
    hip_check(hipGetLastError(), "surrogate step launch");
    std::swap(s->U, s->S);
    s->smax_valid = true;
    if (nsmooth) {
      unsigned long long c = 0;
      hip_check(hipMemcpyAsync(&c, n->nsmooth, sizeof(c), hipMemcpyDeviceToHost, s->st), "copy");
      sync(s);
      *nsmooth = (int64_t)c;
    }
  });
}

}  // extern "C"
