// mm_device.h -- device helpers shared by the HIP kernels: deterministic
// block and group reductions (fixed summation order, no atomics on values).
#pragma once

#include <hip/hip_runtime.h>

namespace mm {

// Sums NV values over the block in a fixed order; the result is valid in
// thread 0.  lds holds 16 doubles per value (blockDim.x <= 1024); its size
// is checked at compile time.
template <int NV, int L>
__device__ inline void block_sum(double (&v)[NV], double (&lds)[L])
{
  static_assert(L >= NV * 16, "block_sum: LDS scratch too small");
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] += __shfl_down(v[i], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; i++) lds[i * 16 + wid] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0.0;
      for (int w = 0; w < nw; w++) s += lds[i * 16 + w];
      v[i] = s;
    }
  }
  __syncthreads();
}

// Run by every thread of a workgroup: sums partial values off .. off+NV-1
// over nslab slabs written by an earlier launch (visible at kernel start, so
// plain cached loads).  Value i of slab b is p0[(off + i) * vstride + b]:
// each value's partials are contiguous, so a workgroup reads nslab * 8 bytes
// per value, not one cache line per slab.  Lane-strided in slab order, then
// block_sum; the total is valid in thread 0, and every workgroup that calls
// it over the same partials obtains the same bits.
template <int NV, int L>
__device__ inline void group_sum_cols(const double *p0, int nslab, long long vstride, int off, double (&out)[NV],
                                      double (&lds)[L])
{
#pragma unroll
  for (int i = 0; i < NV; i++) out[i] = 0.0;
  for (int b = threadIdx.x; b < nslab; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; i++) out[i] += p0[(size_t)(off + i) * vstride + b];
  block_sum<NV>(out, lds);
}

// group_sum_cols four values at a time, the totals written to out[] (LDS,
// valid in every thread after the call); the same bits with fewer live
// registers.
template <int NV, int L>
__device__ inline void group_sum_cols_lds(const double *p0, int nslab, long long vstride, int off, double *out,
                                          double (&lds)[L])
{
#pragma unroll 1
  for (int c = 0; c < NV; c += 4) {
    double v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = 0.0;
    for (int b = threadIdx.x; b < nslab; b += blockDim.x)
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (c + i < NV) v[i] += p0[(size_t)(off + c + i) * vstride + b];
    block_sum<4>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (c + i < NV) out[c + i] = v[i];
  }
  __syncthreads();
}

// The same over row-major partials: value i of slab b at p0[b * stride + i].
template <int NV, int L>
__device__ inline void group_sum_rows(const double *p0, int nslab, int stride, double (&out)[NV],
                                      double (&lds)[L])
{
#pragma unroll
  for (int i = 0; i < NV; i++) out[i] = 0.0;
  for (int b = threadIdx.x; b < nslab; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; i++) out[i] += p0[(size_t)b * stride + i];
  block_sum<NV>(out, lds);
}

}  // namespace mm
