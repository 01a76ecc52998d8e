// Learned face flux: the encoding shared by the data dump, the training kernels and the
// surrogate-assisted step.
//
// A face is described in its normal frame after the floors and the GLM pair
// (glm_physics.jl:317-323): L, R = (rho, un, ua, ub, p, bn*, ba, bb). The network never
// sees raw left/right values (bfloat16 keeps ~3 significant digits, and the quantity of
// interest is a 2% difference); each of the 8 primitives enters as its face mean and its
// jump, both normalised by the local fast-speed scale a = sqrt((gamma p + |B|^2)/rho) of
// the mean state, and jumps are divided by the gate threshold so that gated faces give
// O(1) inputs:
//   x[2v]   = mean_v / scale_v,   x[2v+1] = (R_v - L_v) / (2 scale_v thr)
// with scale = (rho, a, a, a, rho a^2, a sqrt(rho), a sqrt(rho), a sqrt(rho)) of the mean.
// The network predicts the HLLD dissipation relative to the central flux,
//   y_k = (F_HLLD,k - (F(L)_k + F(R)_k)/2) / (thr s_k),
//   s = (rho a, rho a^2, rho a^2, rho a^2, rho a^3, -, a^2 sqrt(rho), a^2 sqrt(rho), -),
// so the flux is F = (F(L)+F(R))/2 + thr s y. The GLM slots (normal B and psi) are not
// learned: they are the exact glm_pair values.
// Gate (the jump detector): |drho|/rho and |dp|/p below thr, the jump of |B| relative to
// its mean at most thr (so that unmagnetised faces pass), and the low-density LLF switch
// not triggered.
#pragma once
#include "mhd_device.hpp"

namespace pmf {

constexpr int SURR_IN = 16, SURR_H = 64, SURR_OUT = 9;

struct SurrScales {
  float s[9];
};

// Encodes a normal-frame face (8 primitives each side, common bn). Returns the gate.
template <class T>
__device__ __forceinline__ bool surr_encode(const T* L, const T* R, T gamma, T thr, T* x, T* s) {
  const T rho = T(0.5) * (L[0] + R[0]);
  const T ir = rcp(rho);
  const T p = T(0.5) * (L[4] + R[4]);
  const T by = T(0.5) * (L[6] + R[6]), bz = T(0.5) * (L[7] + R[7]), bn = L[5];
  const T b2 = bn * bn + by * by + bz * bz;
  const T a2 = (gamma * p + b2) * ir;
  const T ia = rsq(a2);
  const T a = a2 * ia;
  const T sr = dsqrt(rho);
  const T isr = rcp(sr);
  const T sc[8] = {rho, a, a, a, rho * a2, a * sr, a * sr, a * sr};
  const T isc[8] = {ir, ia, ia, ia, ir * ia * ia, ia * isr, ia * isr, ia * isr};
  const T ith = rcp(T(2) * thr);
  (void)sc;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    x[2 * v] = T(0.5) * (L[v] + R[v]) * isc[v];
    x[2 * v + 1] = (R[v] - L[v]) * isc[v] * ith;
  }
  s[0] = thr * rho * a;
  s[1] = s[2] = s[3] = thr * rho * a2;
  s[4] = thr * rho * a2 * a;
  s[5] = T(0);
  s[6] = s[7] = thr * a2 * sr;
  s[8] = T(0);
  const T bL = dsqrt(bn * bn + L[6] * L[6] + L[7] * L[7]);
  const T bR = dsqrt(bn * bn + R[6] * R[6] + R[7] * R[7]);
  const bool smooth = dabs(R[0] - L[0]) < thr * rho && dabs(R[4] - L[4]) < thr * p &&
                      dabs(bR - bL) <= thr * T(0.5) * (bL + bR);  // <=: unmagnetised faces pass
  return smooth;
}

// Central flux (F(L)+F(R))/2 in the normal frame (phys_flux_x, glm_physics.jl:28-34).
template <class T>
__device__ __forceinline__ void surr_central(const T* L, const T* R, T igm1, T* fc) {
  T qL[9] = {L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], T(0)};
  T qR[9] = {R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], T(0)};
  T fL[9], fR[9];
  dflux<0>(qL, fL, igm1);
  dflux<0>(qR, fR, igm1);
#pragma unroll
  for (int k = 0; k < 9; ++k) fc[k] = T(0.5) * (fL[k] + fR[k]);
}

// Normal-frame states after the floors and the GLM pair, and the exact GLM flux slots
// (the first part of riemann, glm_physics.jl:317-324). Returns the low-density switch.
template <class T>
__device__ __forceinline__ bool surr_prepare(T* L, T* R, const StepParams<T>& P, T& fbn, T& fpsi) {
  L[0] = dmax(L[0], P.smallr);
  L[4] = dmax(L[4], P.pfl);
  R[0] = dmax(R[0], P.smallr);
  R[4] = dmax(R[4], P.pfl);
  T bns;
  {
    // the GLM pair exactly as riemann_n evaluates it (no contraction)
#pragma clang fp contract(off)
    bns = T(0.5) * (L[5] + R[5]) - T(0.5) * (R[8] - L[8]) * P.ich;
    fbn = T(0.5) * (L[8] + R[8]) - T(0.5) * P.ch * (R[5] - L[5]);
    fpsi = P.ch2 * bns;
  }
  L[5] = bns;
  R[5] = bns;
  return (P.llf_dmin > T(0) && dmin(L[0], R[0]) < P.llf_dmin) ||
         (P.llf_pmin > T(0) && dmin(L[4], R[4]) < P.llf_pmin);
}

// tanh-form GELU, 0.5 z (1 + tanh(u)) with u = k0 (z + k1 z^3), evaluated as
// z * sigmoid(2u) = z / (1 + exp(-2u)) (one exp2, one reciprocal), and its derivative
// sigma + 2 z sigma (1 - sigma) k0 (1 + 3 k1 z^2); used identically in training and
// inference.
__device__ __forceinline__ float gelu(float z) {
  const float c1 = -2.0f * 1.4426950408889634f * 0.7978845608028654f, c3 = c1 * 0.044715f;
  const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(c3, z * z, c1));  // exp(-2u)
  return z * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float c1 = -2.0f * 1.4426950408889634f * k0, c3 = c1 * k1;
  const float z2 = z * z;
  const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(c3, z2, c1));
  const float sg = __builtin_amdgcn_rcpf(1.0f + e);
  return sg + 2.0f * z * sg * (1.0f - sg) * k0 * (1.0f + 3.0f * k1 * z2);
}

// float -> bfloat16, round to nearest even (the hardware v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ float bf16_round(float f) {
  unsigned u = (unsigned)bf16_bits(f) << 16;
  return __builtin_bit_cast(float, u);
}

}  // namespace pmf

namespace pmf {

// LDS image of the network for wave-level inference: transposed weights W_l^T [out][in]
// in bfloat16 with rows padded by 4 (bank spread, 8-byte aligned MFMA A fragments), the
// output layer padded to 32 rows, then the float32 biases.
constexpr int SIMG_S1 = 16 + 4, SIMG_S2 = 64 + 4;
constexpr int SIMG_W1 = 0;                       // 64 x 20
constexpr int SIMG_W2 = SIMG_W1 + 64 * SIMG_S1;  // 64 x 68
constexpr int SIMG_W3 = SIMG_W2 + 64 * SIMG_S2;  // 64 x 68
constexpr int SIMG_W4 = SIMG_W3 + 64 * SIMG_S2;  // 32 x 68
constexpr int SIMG_NW = SIMG_W4 + 32 * SIMG_S2;  // bf16 entries: 12160
constexpr int SIMG_NB = 64 * 3 + 32;             // float biases (output padded)
constexpr int SIMG_BYTES = SIMG_NW * 2 + SIMG_NB * 4;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef float sf32x16 __attribute__((ext_vector_type(16)));

// Eight bfloat16 weights of one row of the LDS image: four at column c0, four at c1.
__device__ __forceinline__ bf16x8 lds_frag8(const unsigned short* row, int c0, int c1) {
  const u16x4 a = *reinterpret_cast<const u16x4*>(row + c0);
  const u16x4 b = *reinterpret_cast<const u16x4*>(row + c1);
  s16x8 v;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = (short)a[q];
    v[4 + q] = (short)b[q];
  }
  return __builtin_bit_cast(bf16x8, v);
}

// Accumulator initialised with the layer's biases: register r of lane half h holds unit
// 32 mb + 8 (r / 4) + 4 h + r % 4 (the 32x32 MFMA output layout).
__device__ __forceinline__ sf32x16 bias_acc(const float* __restrict__ b, int mb, int lh) {
  sf32x16 d;
#pragma unroll
  for (int r = 0; r < 16; ++r) d[r] = b[32 * mb + 8 * (r / 4) + 4 * lh + r % 4];
  return d;
}

// y[9] = MLP(x[16]) for the 64 faces held by the 64 lanes of the calling wave (one face
// per lane). Must be called by all lanes of the wave. Units are the MFMA rows and faces
// the columns (32 per pass); v_mfma_f32_32x32x16_bf16 contracts 16 units per step. The
// hidden layers' contraction order is permuted so that accumulator registers 8s..8s+7
// of a lane are exactly its B fragment for step s of the next layer (lane half h then
// holds units 16s + 4h + {0..3} and 16s + 8 + 4h + {0..3}); the weight fragments read
// the same units. Activations never leave registers.
// npass = 1 evaluates only the faces of lanes 0-31 (results in lanes 0-31).
__device__ inline __attribute__((noinline)) void mlp_wave(const float* x, float* y,
                                                        const unsigned short* __restrict__ sw,
                                                        const float* __restrict__ sb, int npass = 2) {
  const int lane = threadIdx.x & 63, li = lane & 31, lh = lane >> 5;
#pragma unroll 1
  for (int pass = 0; pass < npass; ++pass) {
    const int src = pass * 32 + li;
    // layer 1: inputs 8h..8h+7 of face (pass, li) in lane half h
    bf16x8 bin;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float lo = __shfl(x[q], src, 64), hi = __shfl(x[8 + q], src, 64);
      bin[q] = (__bf16)(lh ? hi : lo);
    }
    sf32x16 d[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
      d[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
          lds_frag8(sw + SIMG_W1 + (mb * 32 + li) * SIMG_S1, 8 * lh, 8 * lh + 4), bin, bias_acc(sb, mb, lh), 0, 0, 0);
#pragma unroll 1
    for (int layer = 1; layer < 4; ++layer) {
      bf16x8 b[4];
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int q = 0; q < 8; ++q) b[st][q] = (__bf16)gelu(d[st >> 1][8 * (st & 1) + q]);
      const int off = layer == 1 ? SIMG_W2 : layer == 2 ? SIMG_W3 : SIMG_W4;
      const int nmb = layer == 3 ? 1 : 2;
      const float* bias = sb + 64 * layer;
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        if (mb >= nmb) break;
        sf32x16 acc = bias_acc(bias, mb, lh);
        const unsigned short* row = sw + off + (mb * 32 + li) * SIMG_S2;
#pragma unroll
        for (int st = 0; st < 4; ++st)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_frag8(row, 16 * st + 4 * lh, 16 * st + 8 + 4 * lh), b[st],
                                                         acc, 0, 0, 0);
        d[mb] = acc;
      }
    }
    // outputs o of face column j: o = 0..3 in lane j (registers 0..3), o = 4..7 in lane
    // j+32 (registers 0..3), o = 8 in lane j (register 4)
    float o[9];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      o[q] = __shfl(d[0][q], li, 64);
      o[4 + q] = __shfl(d[0][q], li + 32, 64);
    }
    o[8] = __shfl(d[0][4], li, 64);
    if (lh == pass) {
#pragma unroll
      for (int k = 0; k < 9; ++k) y[k] = o[k];
    }
  }
}

}  // namespace pmf


namespace pmf {

// The gated face flux in two halves around a batched network pass (the surrogate-assisted
// fused step queues the smooth faces of a whole tile plane and evaluates them in dense
// batches of 64, so waves whose faces are mostly rough do not run the network).
// surr_face_begin: lab-frame states in; in the normal frame, f is the exact flux of a
// rough face (HLLD, or LLF under the low-density switch) or the central flux of a smooth
// one with the exact GLM slots, s the output scales and x the 16 network inputs. Returns
// whether the face is smooth. Lanes with active false compute nothing of use.
template <int D>
__device__ __forceinline__ bool surr_face_begin(const float* Lq, const float* Rq, const StepParams<float>& P,
                                                float thr, bool active, float* f, float* s, float* x) {
  constexpr int N = D, A = (D + 1) % 3, B = (D + 2) % 3;
  float L[9] = {Lq[0], Lq[1 + N], Lq[1 + A], Lq[1 + B], Lq[4], Lq[5 + N], Lq[5 + A], Lq[5 + B], Lq[8]};
  float R[9] = {Rq[0], Rq[1 + N], Rq[1 + A], Rq[1 + B], Rq[4], Rq[5 + N], Rq[5 + A], Rq[5 + B], Rq[8]};
  float fbn, fpsi;
  const bool uself = surr_prepare(L, R, P, fbn, fpsi);
  const bool smooth = surr_encode(L, R, P.gamma, thr, x, s) && !uself && active;
  const bool rough = active && !smooth;
  if (rough) {
    if (uself) llf_n(L, R, P.ch, fbn, fpsi, f, P);
    else hlld_n(L, R, L[5], fbn, fpsi, f, P);
  }
  if (smooth) {
    surr_central(L, R, P.igm1, f);
    f[5] = fbn;
    f[8] = fpsi;
  }
  return smooth;
}

// surr_face_end: adds the learned dissipation y (smooth faces) and rotates to the lab frame.
template <int D>
__device__ __forceinline__ void surr_face_end(float* f, const float* s, const float* y, bool smooth, float* F) {
  constexpr int N = D, A = (D + 1) % 3, B = (D + 2) % 3;
  if (smooth) {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (k != 5 && k != 8) f[k] = f[k] + s[k] * y[k];
  }
  F[0] = f[0];
  F[1 + N] = f[1];
  F[1 + A] = f[2];
  F[1 + B] = f[3];
  F[4] = f[4];
  F[5 + N] = f[5];
  F[5 + A] = f[6];
  F[5 + B] = f[7];
  F[8] = f[8];
}

// LDS queue of the smooth faces of one tile plane: one 32-byte slot per face holding its
// 16 network inputs as bfloat16 (the precision the network reads them in), which the
// network pass overwrites with the seven learned outputs in float32 (the GLM slots 5 and
// 8 are exact and not learned).
constexpr int SURR_QOUT = 7;
__device__ __forceinline__ constexpr int surr_qk(int j) { return j < 5 ? j : j + 1; }  // output slot j -> k

template <int CAP>
struct SurrQueue {
  unsigned n;
  unsigned pad[3];
  uint4 slot[CAP * 2];  // 32 bytes per face
};

// Appends the calling lanes' faces that want a slot (one LDS atomic per wave); returns
// the slot, or -1.
template <int CAP>
__device__ __forceinline__ int surr_enqueue(bool want, const float* x, SurrQueue<CAP>& Q) {
  const unsigned long long b = __ballot(want);
  if (b == 0ull) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(b);
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(&Q.n, (unsigned)__popcll(b));
  base = __shfl(base, leader, 64);
  if (!want) return -1;
  const int slot = int(base) + __popcll(b & ((1ull << lane) - 1ull));
  unsigned w[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = (unsigned)bf16_bits(x[2 * k]) | ((unsigned)bf16_bits(x[2 * k + 1]) << 16);
  Q.slot[2 * slot] = make_uint4(w[0], w[1], w[2], w[3]);
  Q.slot[2 * slot + 1] = make_uint4(w[4], w[5], w[6], w[7]);
  return slot;
}

// One batch of up to 32 queued faces (slots b .. b+31 below nq, one MFMA pass), evaluated
// by the calling wave (lanes 0-31 hold the faces); all lanes of the wave must call it.
template <int CAP>
__device__ __forceinline__ void surr_batch(int b, int nq, SurrQueue<CAP>& Q, const unsigned short* __restrict__ sw,
                                           const float* __restrict__ sb) {
  const int lane = threadIdx.x & 63;
  const int slot = b + lane;
  const bool valid = lane < 32 && slot < nq;
  float x[SURR_IN];
  if (valid) {
    const uint4 a = Q.slot[2 * slot];
    const uint4 c = Q.slot[2 * slot + 1];
    const unsigned w[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      x[2 * k] = __builtin_bit_cast(float, w[k] << 16);
      x[2 * k + 1] = __builtin_bit_cast(float, w[k] & 0xffff0000u);
    }
  } else {
#pragma unroll
    for (int k = 0; k < SURR_IN; ++k) x[k] = 0.f;
  }
  float y[SURR_OUT];
  mlp_wave(x, y, sw, sb, 1);
  // the slot's inputs were read above by this lane, the only one that touches the slot
  if (valid) {
    float* out = reinterpret_cast<float*>(&Q.slot[2 * slot]);
#pragma unroll
    for (int j = 0; j < SURR_QOUT; ++j) out[j] = y[surr_qk(j)];
  }
}

// The learned outputs of queue slot q (zeros for q < 0), after the network pass.
template <int CAP>
__device__ __forceinline__ void surr_result(const SurrQueue<CAP>& Q, int q, float* y) {
#pragma unroll
  for (int k = 0; k < SURR_OUT; ++k) y[k] = 0.f;
  if (q >= 0) {
    const float* in = reinterpret_cast<const float*>(&Q.slot[2 * q]);
#pragma unroll
    for (int j = 0; j < SURR_QOUT; ++j) y[surr_qk(j)] = in[j];
  }
}

}  // namespace pmf
