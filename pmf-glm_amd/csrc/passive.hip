// Passive species transport after the MHD step (lib/MHDKernels/src/passive_transport.jl):
// conserved species densities to bounded mass fractions, and the semi-Lagrangian
// backtrace of fractions (or of an unconstrained scalar) with the post-step cell velocity,
// sampled trilinearly on the periodic cubic grid and restored to mass densities with the
// new gas density.
#include <limits>

#include "launch.hpp"

namespace pmf {

template <class T> __device__ __forceinline__ T clampT(T x, T lo, T hi) { return x < lo ? lo : (x > hi ? hi : x); }

// passive_transport.jl:5-25
template <class T>
__global__ __launch_bounds__(256) void k_species_to_fractions(T* __restrict__ s1, T* __restrict__ s2,
                                                              const T* __restrict__ rho, int64_t nc, T maxt) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  const T d = dmax(rho[c], std::numeric_limits<T>::epsilon());
  if (s2) {
    const T q2 = clampT(s2[c] / d, T(0), maxt);
    const T q1 = clampT(s1[c] / d, T(0), maxt - q2);
    s1[c] = q1;
    s2[c] = q2;
  } else {
    s1[c] = clampT(s1[c] / d, T(0), maxt);
  }
}

// passive_transport.jl:27-32 (0-based, non-negative modulo)
__device__ __forceinline__ int64_t passive_index(int i, int j, int k, int n) {
  i %= n;
  j %= n;
  k %= n;
  i += i < 0 ? n : 0;
  j += j < 0 ? n : 0;
  k += k < 0 ? n : 0;
  return (int64_t)i + (int64_t)n * ((int64_t)j + (int64_t)n * k);
}

// passive_transport.jl:40-54: muladd(f, r - l, l) along x, then y, then z, as one fused
// multiply-add in the field's own precision
__device__ __forceinline__ float lerp(float l, float r, float f) { return __builtin_fmaf(f, r - l, l); }
__device__ __forceinline__ double lerp(double l, double r, double f) { return __builtin_fma(f, r - l, l); }
template <class T> __device__ __forceinline__ T trilinear(const T* __restrict__ a, int i0, int j0, int k0, int n, T fx,
                                                          T fy, T fz) {
  const T p000 = a[passive_index(i0, j0, k0, n)], p100 = a[passive_index(i0 + 1, j0, k0, n)];
  const T p010 = a[passive_index(i0, j0 + 1, k0, n)], p110 = a[passive_index(i0 + 1, j0 + 1, k0, n)];
  const T p001 = a[passive_index(i0, j0, k0 + 1, n)], p101 = a[passive_index(i0 + 1, j0, k0 + 1, n)];
  const T p011 = a[passive_index(i0, j0 + 1, k0 + 1, n)], p111 = a[passive_index(i0 + 1, j0 + 1, k0 + 1, n)];
  const T x00 = lerp(p000, p100, fx), x10 = lerp(p010, p110, fx);
  const T x01 = lerp(p001, p101, fx), x11 = lerp(p011, p111, fx);
  return lerp(lerp(x00, x10, fy), lerp(x01, x11, fy), fz);
}

// passive_transport.jl:78-190. mode 0: one fraction, 1: two fractions (the second
// bounds the first), 2: unconstrained scalar.
template <class T>
__global__ __launch_bounds__(256) void k_backtrace(T* __restrict__ o1, T* __restrict__ o2, const T* __restrict__ a1,
                                                   const T* __restrict__ a2, const T* __restrict__ rho,
                                                   const T* __restrict__ mx, const T* __restrict__ my,
                                                   const T* __restrict__ mz, int n, T dtdx, T maxt, int mode) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nc = (int64_t)n * n * n;
  if (c >= nc) return;
  const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / ((int64_t)n * n));
  const T d = dmax(rho[c], std::numeric_limits<T>::epsilon());
  const T x = T(i) - dtdx * mx[c] / d;
  const T y = T(j) - dtdx * my[c] / d;
  const T z = T(k) - dtdx * mz[c] / d;
  const int i0 = (int)__builtin_floor(x), j0 = (int)__builtin_floor(y), k0 = (int)__builtin_floor(z);
  const T fx = x - T(i0), fy = y - T(j0), fz = z - T(k0);
  const T s1 = trilinear(a1, i0, j0, k0, n, fx, fy, fz);
  if (mode == 2) {
    o1[c] = s1;
  } else if (mode == 1) {
    const T s2 = trilinear(a2, i0, j0, k0, n, fx, fy, fz);
    const T q2 = clampT(s2, T(0), maxt);
    const T q1 = clampT(s1, T(0), maxt - q2);
    o1[c] = d * q1;
    o2[c] = d * q2;
  } else {
    o1[c] = d * clampT(s1, T(0), maxt);
  }
}

template <class T>
void launch_species_to_fractions(T* s1, T* s2, const T* rho, int64_t nc, T maxt, hipStream_t st) {
  hipLaunchKernelGGL(k_species_to_fractions<T>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, s1, s2, rho, nc,
                     maxt);
}

template <class T>
void launch_backtrace(T* o1, T* o2, const T* a1, const T* a2, Fields<const T> U, int n, T dtdx, T maxt, int mode,
                      hipStream_t st) {
  const int64_t nc = (int64_t)n * n * n;
  hipLaunchKernelGGL(k_backtrace<T>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, o1, o2, a1, a2, U.v[0],
                     U.v[1], U.v[2], U.v[3], n, dtdx, maxt, mode);
}

template void launch_species_to_fractions<float>(float*, float*, const float*, int64_t, float, hipStream_t);
template void launch_species_to_fractions<double>(double*, double*, const double*, int64_t, double, hipStream_t);
template void launch_backtrace<float>(float*, float*, const float*, const float*, Fields<const float>, int, float,
                                      float, int, hipStream_t);
template void launch_backtrace<double>(double*, double*, const double*, const double*, Fields<const double>, int,
                                       double, double, int, hipStream_t);

}  // namespace pmf
