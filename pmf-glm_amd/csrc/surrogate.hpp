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
  T bns = T(0.5) * (L[5] + R[5]) - T(0.5) * (R[8] - L[8]) * P.ich;
  fbn = T(0.5) * (L[8] + R[8]) - T(0.5) * P.ch * (R[5] - L[5]);
  fpsi = P.ch2 * bns;
  L[5] = bns;
  R[5] = bns;
  return (P.llf_dmin > T(0) && dmin(L[0], R[0]) < P.llf_dmin) ||
         (P.llf_pmin > T(0) && dmin(L[4], R[4]) < P.llf_pmin);
}

// tanh-form GELU and its derivative (used identically in training and inference).
__device__ __forceinline__ float gelu(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  float u = k0 * (z + k1 * z * z * z);
  float e = __builtin_amdgcn_exp2f(2.0f * 1.4426950408889634f * u);  // exp(2u)
  float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);            // tanh(u)
  return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  float u = k0 * (z + k1 * z * z * z);
  float e = __builtin_amdgcn_exp2f(2.0f * 1.4426950408889634f * u);
  float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * k0 * (1.0f + 3.0f * k1 * z * z);
}

// float -> bfloat16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
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

typedef short sbf16x4 __attribute__((ext_vector_type(4)));
typedef float sf32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ sbf16x4 lds_frag(const unsigned short* base) {
  return *reinterpret_cast<const sbf16x4*>(base);
}

// y[9] = MLP(x[16]) for the 64 faces held by the 64 lanes of the calling wave (one face
// per lane). Must be called by all lanes of the wave. Units are the MFMA rows and faces
// the columns (32 per pass), so a layer's 32x32 accumulator registers 4s..4s+3 are
// exactly the next layer's B fragment for k-step s: activations stay in registers.
__device__ __forceinline__ void mlp_wave(const float* x, float* y, const unsigned short* __restrict__ sw,
                                         const float* __restrict__ sb) {
  const int lane = threadIdx.x & 63, li = lane & 31, lh = lane >> 5;
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    const int src = pass * 32 + li;
    sbf16x4 bin[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float lo = __shfl(x[8 * ks + q], src, 64), hi = __shfl(x[8 * ks + 4 + q], src, 64);
        bin[ks][q] = (short)bf16_bits(lh ? hi : lo);
      }
    sf32x16 d[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
      for (int r = 0; r < 16; ++r) d[mb][r] = 0.f;
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
        d[mb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
            lds_frag(sw + SIMG_W1 + (mb * 32 + li) * SIMG_S1 + 8 * ks + 4 * lh), bin[ks], d[mb], 0, 0, 0);
    }
#pragma unroll 1
    for (int layer = 1; layer < 4; ++layer) {
      sbf16x4 b[8];
      const float* bias = sb + 64 * (layer - 1);
#pragma unroll
      for (int ks = 0; ks < 8; ++ks)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 4 * (ks & 3) + q;
          const int u = (ks >> 2) * 32 + 8 * (ks & 3) + 4 * lh + q;
          b[ks][q] = (short)bf16_bits(gelu(d[ks >> 2][r] + bias[u]));
        }
      const int off = layer == 1 ? SIMG_W2 : layer == 2 ? SIMG_W3 : SIMG_W4;
      const int nmb = layer == 3 ? 1 : 2;
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        if (mb >= nmb) break;
#pragma unroll
        for (int r = 0; r < 16; ++r) d[mb][r] = 0.f;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks)
          d[mb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
              lds_frag(sw + off + (mb * 32 + li) * SIMG_S2 + 8 * ks + 4 * lh), b[ks], d[mb], 0, 0, 0);
      }
    }
    // outputs o of face column j: o = 0..3 in lane j (registers 0..3), o = 4..7 in lane
    // j+32 (registers 0..3), o = 8 in lane j (register 4)
    const float* b4 = sb + 192;
    float o[9];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      o[q] = __shfl(d[0][q], li, 64);
      o[4 + q] = __shfl(d[0][q], li + 32, 64);
    }
    o[8] = __shfl(d[0][4], li, 64);
    if (lh == pass) {
#pragma unroll
      for (int k = 0; k < 9; ++k) y[k] = o[k] + b4[k];
    }
  }
}

}  // namespace pmf

namespace pmf {

// Gated face flux for the surrogate-assisted step: lab-frame states in, lab-frame flux
// out. Must be called by all lanes of a wave (active marks lanes that own a face; the
// others pass any finite state). Rough or inactive-adjacent faces take HLLD (LLF under
// the low-density switch), smooth ones the central flux plus the learned dissipation.
// Returns whether this lane's face used the network.
template <int D>
__device__ __forceinline__ bool surr_face(const float* Lq, const float* Rq, float* F, const StepParams<float>& P,
                                          float thr, bool active, const unsigned short* sw, const float* sb) {
  constexpr int N = D, A = (D + 1) % 3, B = (D + 2) % 3;
  float L[9] = {Lq[0], Lq[1 + N], Lq[1 + A], Lq[1 + B], Lq[4], Lq[5 + N], Lq[5 + A], Lq[5 + B], Lq[8]};
  float R[9] = {Rq[0], Rq[1 + N], Rq[1 + A], Rq[1 + B], Rq[4], Rq[5 + N], Rq[5 + A], Rq[5 + B], Rq[8]};
  float fbn, fpsi;
  const bool uself = surr_prepare(L, R, P, fbn, fpsi);
  float x[16], s[9], f[9];
  const bool smooth = surr_encode(L, R, P.gamma, thr, x, s) && !uself && active;
  const bool rough = active && !smooth;
  if (__any(rough)) {
    if (rough) {
      if (uself) llf_n(L, R, P.ch, fbn, fpsi, f, P);
      else hlld_n(L, R, L[5], fbn, fpsi, f, P);
    }
  }
  if (__any(smooth)) {
    float y[9];
    mlp_wave(x, y, sw, sb);
    if (smooth) {
      float fc[9];
      surr_central(L, R, P.igm1, fc);
#pragma unroll
      for (int k = 0; k < 9; ++k) f[k] = fc[k] + s[k] * y[k];
      f[5] = fbn;
      f[8] = fpsi;
    }
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
  return smooth;
}

}  // namespace pmf
