// mm_device.h -- device helpers shared by the HIP kernels: deterministic
// block reductions and the "last block of a group" completion pattern used
// to finish a dot product without an extra kernel launch or atomics on the
// values themselves.
#pragma once

#include <hip/hip_runtime.h>

namespace mm {

// Device-coherent load (bypasses the non-coherent per-CU vector cache) for
// values other blocks wrote during the same launch.
__device__ inline double load_coherent(const double *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Called by every thread of every block of a group of `nblocks` blocks after
// the block's partial results are stored.  Returns true in exactly one block,
// the last to arrive, once all partials of the group are visible.  That block
// resets the counter for the next launch.  No block waits on another.
// Every block publishes with an agent-scope release only (an L2 write-back on
// a multi-XCD part); the acquire, which invalidates the XCD's L2, is paid
// once, by the last block.  A full __threadfence() in every block would
// invalidate L2 thousands of times per launch.
__device__ inline bool arrive_last(int *counter, int nblocks)
{
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const int prev = atomicAdd(counter, 1);
    s_last = (prev == nblocks - 1);
    if (s_last) {
      atomicExch(counter, 0);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }
  __syncthreads();
  return s_last != 0;
}

// Run by every thread of the last workgroup of a group: sums NV consecutive
// partials (offset off) over the group's nslab partial rows (row stride
// `stride`) in a fixed order (lane-strided, then block_sum).  The total is
// valid in thread 0.
template <int NV, int L>
__device__ inline void group_reduce(const double *p0, int nslab, int stride, int off, double (&out)[NV],
                                    double (&lds)[L])
{
#pragma unroll
  for (int i = 0; i < NV; i++) out[i] = 0.0;
  for (int b = threadIdx.x; b < nslab; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; i++) out[i] += load_coherent(&p0[(size_t)b * stride + off + i]);
  block_sum<NV>(out, lds);
}

// Like group_reduce, for partials written by an earlier launch (visible at
// kernel start, so plain cached loads).  Every workgroup that calls it over
// the same partials obtains the same bits.
template <int NV, int L>
__device__ inline void group_sum(const double *p0, int nslab, int stride, int off, double (&out)[NV],
                                 double (&lds)[L])
{
#pragma unroll
  for (int i = 0; i < NV; i++) out[i] = 0.0;
  for (int b = threadIdx.x; b < nslab; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < NV; i++) out[i] += p0[(size_t)b * stride + off + i];
  block_sum<NV>(out, lds);
}

}  // namespace mm
