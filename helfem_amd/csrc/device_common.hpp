// Device helpers shared by the atomic and diatomic kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

namespace hfg {
namespace dev {

constexpr int NT = 256;  // threads per workgroup (4 wave64)

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Reduce two values over the 256-thread block; result valid in thread 0.
__device__ inline void block_sum2(double &x, double &y, double *red) {
  x = wave_sum(x);
  y = wave_sum(y);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[w] = x;
    red[4 + w] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    x = red[0] + red[1] + red[2] + red[3];
    y = red[4] + red[5] + red[6] + red[7];
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));

// D(16x16) += X(16x16) * Y(16x16), one wave, four v_mfma_f64_16x16x4f64.
// Lane l (g = l/16, c = l%16) supplies X(c, k) and Y(k, c) for the inner
// index k = g + 4s of step s, and holds the result D(g + 4r, c) in component
// r (the layout measured on gfx950).  With this choice of k a result tile is
// already in Y-operand layout, so products chain without a shuffle.
__device__ inline v4d mfma_16x16x16(const double xs[4], const double ys[4], v4d d) {
#pragma unroll
  for (int s = 0; s < 4; ++s) d = __builtin_amdgcn_mfma_f64_16x16x4f64(xs[s], ys[s], d, 0, 0, 0);
  return d;
}

__device__ inline v4d mfma_16x16x16(const double xs[4], v4d y, v4d d) {
  const double ys[4] = {y[0], y[1], y[2], y[3]};
  return mfma_16x16x16(xs, ys, d);
}

typedef float v4f __attribute__((ext_vector_type(4)));

// Operand/result layout of the 16x16x4 matrix instructions, per precision.
// Both take X(c, k) and Y(k, c) from lane l (g = l/16, c = l%16) with the
// instruction's inner index g; the result row held in accumulator component
// r differs (measured on gfx950 with one-hot probes):
//   v_mfma_f64_16x16x4f64: D(g + 4r, c)      v_mfma_f32_16x16x4f32: D(4g + r, c)
// Choosing the global inner index of step s as row(g, s) makes a result tile
// the next product's Y operand without a shuffle, in either precision.
template <typename T> struct Mma;
template <> struct Mma<double> {
  typedef v4d acc_t;
  __device__ static int row(int g, int s) { return g + 4 * s; }
  __device__ static acc_t step(double x, double y, acc_t d) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, d, 0, 0, 0);
  }
};
template <> struct Mma<float> {
  typedef v4f acc_t;
  __device__ static int row(int g, int s) { return 4 * g + s; }
  __device__ static acc_t step(float x, float y, acc_t d) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, d, 0, 0, 0);
  }
};

// D(16x16) += X Y over the 16-long inner index, four instructions.
template <typename T>
__device__ inline typename Mma<T>::acc_t mma16(const T xs[4], const T ys[4], typename Mma<T>::acc_t d) {
#pragma unroll
  for (int s = 0; s < 4; ++s) d = Mma<T>::step(xs[s], ys[s], d);
  return d;
}
template <typename T>
__device__ inline typename Mma<T>::acc_t mma16(const T xs[4], typename Mma<T>::acc_t y,
                                               typename Mma<T>::acc_t d) {
#pragma unroll
  for (int s = 0; s < 4; ++s) d = Mma<T>::step(xs[s], y[s], d);
  return d;
}

// Index s of the system owning workgroup b of a launch that concatenates
// the grids of n systems: the largest s with start[s] <= b (systems with no
// workgroups have start[s] == start[s+1] and are skipped).
__device__ inline int batch_find(const int *start, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  // the same for every lane: kept in a scalar register, so the system's
  // argument struct is read with scalar loads
  return __builtin_amdgcn_readfirstlane(lo);
}

}  // namespace dev
}  // namespace hfg
