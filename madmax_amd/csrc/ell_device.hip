// ell_device.hip -- device mirror of the reference ELL matrix and the GPU
// compute entry points of include/ell.h (ell_mvp, ell_solve_jacobi,
// ell_solve_pcg).
//
// HBM layout: rows are grouped into slabs of S rows (S a multiple of 64);
// inside a slab the slots are column-major, element (row, j) at
// slab*S*nnz + j*S + row%S, so a wavefront reading slot j of 64 consecutive
// rows issues one coalesced 256-byte (int32) / 512-byte (FP64) access.
// Padding rows of the last slab hold cols = -1, vals = 0.
//
// Bit-exactness: the SpMV accumulates each row in ascending slot order,
// stops at the first -1 and keeps multiply and add separate, which is the
// arithmetic of src/ell.c:127-141 (the reference is built without FMA).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "ell.h"
#include "fe.h"
#include "mm_device.h"
#include "mm_internal.h"

using namespace mm;

namespace {

struct IterState {
  double bnorm, rnorm, rz, alpha, beta, err;
  int its, done, max_its, pbuf;
  double rtol, min_err;
  int pad0, pad1, pad2;  // k_pcg_fused: fault flag, and the counts a failed barrier wait saw
  int fresh;             // PCG: 1 from the start until the first update's partials exist
};

// Dot products: every workgroup of a big kernel writes its partial sums
// (part[slab * 4 + i]).  Jacobi and the PCG start sum them in a one-workgroup
// kernel (k_reduce) in slab order; in the PCG iteration every workgroup of
// the consuming kernel sums them itself, in the same order, so all hold the
// same scalars.  Kernel boundaries order producer and consumer, so no
// workgroup fences for another.
enum RedMode : int { RED_JAC = 0, RED_INIT = 1, RED_PQ = 2, RED_RR = 3 };

// The scalar steps of Jacobi-PCG, shared by every kernel that takes them so
// that they round alike.  After q = A p: alpha from p.q.  After the update:
// count the iteration, test ||r|| <= rtol ||b|| or the cap, else beta and
// the new r.z.
__device__ inline double pcg_alpha(double rz, double pq) { return rz / pq; }
__device__ inline void pcg_after_update(double rr, double rzn, double rtol, double bnorm, int max_its, int &its,
                                        double &rnorm, double &rz, double &beta, int &done)
{
  its += 1;
  rnorm = sqrt(rr);
  if (rnorm <= rtol * bnorm || its >= max_its) {
    done = 1;
  } else {
    beta = rzn / rz;
    rz = rzn;
  }
}

// Row `lane` of the slab times x: ascending slots, stop at the first -1,
// separate multiply and add.  Slots are fetched kU at a time (column ids and
// values in one round trip, the x entries in a second) and then accumulated
// in slot order, so the loads of a batch overlap instead of waiting on each
// other.  Column ids are -1 or in [0, ncol) (checked on upload), so the
// entries fetched past a row's first -1 are in bounds and are never added.
constexpr int kU = 8;
// Matrices of at most this many slabs sum the PCG partials in every
// workgroup (two launches per iteration); larger ones reduce them in
// k_reduce launches (four).  On one MI355X: 350 slabs (30^3 RVE) 0.042 ms
// against 0.045 ms per iteration; 12,070 slabs (100^3) 0.89 against 0.73.
constexpr int kSelfSumSlabs = 1024;
__device__ inline double row_dot_exact(const int *__restrict__ cols, const double *__restrict__ vals,
                                       long long slab_base, int S, int nnz, int lane,
                                       const double *__restrict__ x)
{
#pragma clang fp contract(off)
  double acc = 0.0;
  for (int j0 = 0; j0 < nnz; j0 += kU) {
    int c[kU];
    double v[kU], xv[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const bool in = j0 + u < nnz;
      const long long e = slab_base + (long long)(j0 + u) * S + lane;
      c[u] = in ? __builtin_nontemporal_load(&cols[e]) : -1;
      v[u] = in ? __builtin_nontemporal_load(&vals[e]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) xv[u] = (c[u] >= 0) ? x[c[u]] : 0.0;
    bool stop = false;
#pragma unroll
    for (int u = 0; u < kU; u++) {
      if (c[u] == -1) {
        stop = true;
        break;
      }
      acc = acc + v[u] * xv[u];
    }
    if (stop) break;
  }
  return acc;
}

// y = A x, one lane per row (src/ell.c:127-141)
__global__ void k_spmv64_exact(const int *__restrict__ cols, const double *__restrict__ vals,
                               int nrow, int nnz, int S, const double *__restrict__ x,
                               double *__restrict__ y)
{
  const int lane = threadIdx.x;
  const long long row = (long long)blockIdx.x * S + lane;
  if (row >= nrow) return;
  const long long base = (long long)blockIdx.x * S * nnz;
  y[row] = row_dot_exact(cols, vals, base, S, nnz, lane, x);
}

// The one-workgroup reduction and the scalar steps of both solvers.
__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ part, int nslab, int mode,
                                                IterState *st)
{
  __shared__ double lds[48];
  if (st->done) return;
  if (mode == RED_JAC && st->its >= st->max_its) {  // loop condition of the reference
    if (threadIdx.x == 0) st->done = 1;
    return;
  }
  double s[3];
  group_sum_rows<3>(part, nslab, 4, s, lds);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case RED_JAC:  // src/ell.c:160-180: err = ||A x - b||; stop, or update and count
      st->err = sqrt(s[0]);
      if (st->err < st->min_err) st->done = 1;
      else st->its += 1;
      break;
    case RED_INIT:
      st->bnorm = sqrt(s[0]);
      st->rnorm = sqrt(s[1]);
      st->rz = s[2];
      st->beta = 0.0;
      st->its = 0;
      st->pbuf = 0;
      st->fresh = 1;
      st->done = (st->bnorm == 0.0 || st->rnorm <= st->rtol * st->bnorm) ? 1 : 0;
      break;
    case RED_PQ:
      st->alpha = pcg_alpha(st->rz, s[0]);
      st->pbuf ^= 1;
      break;
    default:  // RED_RR: r.r and r.z are k_pcg_update's part[.][1..2]
      pcg_after_update(s[1], s[2], st->rtol, st->bnorm, st->max_its, st->its, st->rnorm, st->rz, st->beta, st->done);
  }
}

// ---- stationary Jacobi (src/ell.c:143-184) --------------------------------
// k_jac_res: r = A x - b and the partial ||r||^2 (part[.][0]).
__global__ void k_jac_res(const int *__restrict__ cols, const double *__restrict__ vals, int nrow,
                          int nnz, int S, const double *__restrict__ x, const double *__restrict__ b,
                          double *__restrict__ r, double *__restrict__ part, const IterState *st)
{
#pragma clang fp contract(off)
  __shared__ double lds[16];
  if (st->done || st->its >= st->max_its) return;
  const int lane = threadIdx.x;
  const long long row = (long long)blockIdx.x * S + lane;
  double v[1] = {0.0};
  if (row < nrow) {
    const long long base = (long long)blockIdx.x * S * nnz;
    double ri = row_dot_exact(cols, vals, base, S, nnz, lane, x);
    ri = ri - b[row];
    r[row] = ri;
    v[0] = ri * ri;
  }
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 4] = v[0];
    part[blockIdx.x * 4 + 1] = 0.0;
    part[blockIdx.x * 4 + 2] = 0.0;
  }
}

// k_jac_upd: x -= dinv * r
__global__ void k_jac_upd(int nrow, int S, const double *__restrict__ dinv,
                          const double *__restrict__ r, double *__restrict__ x, const IterState *st)
{
#pragma clang fp contract(off)
  if (st->done) return;
  const long long row = (long long)blockIdx.x * S + threadIdx.x;
  if (row >= nrow) return;
  const double e = dinv[row] * r[row];
  x[row] = x[row] - e;
}

// ---- Jacobi-preconditioned CG in FP64 ---------------------------------------
// init: r = b - A x, z = dinv r; partials ||b||^2, ||r||^2, r.z
__global__ void k_pcg_init(const int *__restrict__ cols, const double *__restrict__ vals, int nrow,
                           int nnz, int S, const double *__restrict__ x, const double *__restrict__ b,
                           const double *__restrict__ dinv, double *__restrict__ r,
                           double2 *__restrict__ zp0, double2 *__restrict__ zp1, double *__restrict__ part)
{
  __shared__ double lds[48];
  const int lane = threadIdx.x;
  const long long row = (long long)blockIdx.x * S + lane;
  double v[3] = {0.0, 0.0, 0.0};
  if (row < nrow) {
    const long long base = (long long)blockIdx.x * S * nnz;
    const double ax = row_dot_exact(cols, vals, base, S, nnz, lane, x);
    const double bi = b[row];
    const double ri = bi - ax;
    const double zi = dinv[row] * ri;
    r[row] = ri;
    zp0[row] = make_double2(zi, 0.0);
    zp1[row] = make_double2(0.0, 0.0);
    v[0] = bi * bi;
    v[1] = ri * ri;
    v[2] = ri * zi;
  }
  block_sum<3>(v, lds);
  if (threadIdx.x == 0)
    for (int i = 0; i < 3; i++) part[blockIdx.x * 4 + i] = v[i];
}

// z and p are stored as pairs {z, p} per row, in two buffers that the
// iterations alternate between (the state's pbuf): q = A p reads the current
// buffer's pairs (one 16-byte gather per column) and writes the new p into
// the other buffer, whose z the update then writes.

// Row `row` (lane `lane` of slab `slab`) of q = A p with p = z + beta p_old
// formed on the fly from the current pairs zc, both for the own row, whose p
// is stored in zn, and for every gathered column; returns q_row * p_row.
// Shared by the per-step kernels and the fused kernel, so both do the same
// arithmetic.
__device__ inline double pcg_q_row(const int *__restrict__ cols, const double *__restrict__ vals, int nnz, int S,
                                   int slab, int lane, long long row, const double2 *__restrict__ zc, double beta,
                                   double2 *__restrict__ zn, double *__restrict__ q)
{
  const long long base = (long long)slab * S * nnz;
  double acc = 0.0;
  // kU slots per round trip, as in row_dot_exact; accumulated in slot order
  for (int j0 = 0; j0 < nnz; j0 += kU) {
    int c[kU];
    double v[kU], t[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const bool in = j0 + u < nnz;
      const long long e = base + (long long)(j0 + u) * S + lane;
      c[u] = in ? __builtin_nontemporal_load(&cols[e]) : -1;
      v[u] = in ? __builtin_nontemporal_load(&vals[e]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const double2 w = (c[u] >= 0) ? zc[c[u]] : make_double2(0.0, 0.0);
      t[u] = w.x + beta * w.y;
    }
    bool stop = false;
#pragma unroll
    for (int u = 0; u < kU; u++) {
      if (c[u] == -1) {
        stop = true;
        break;
      }
      acc += v[u] * t[u];
    }
    if (stop) break;
  }
  const double2 w = zc[row];
  const double pi = w.x + beta * w.y;
  zn[row].y = pi;
  q[row] = acc;
  return pi * acc;
}

// Row `row` of x += alpha p, r -= alpha q, z = dinv r, with p and the new z
// in the pairs zp; v = {r.r, r.z} terms.
__device__ inline void pcg_update_row(long long row, double alpha, const double *__restrict__ dinv,
                                      double2 *__restrict__ zp, const double *__restrict__ q,
                                      double *__restrict__ x, double *__restrict__ r, double (&v)[2])
{
  const double xi = x[row] + alpha * zp[row].y;
  const double ri = r[row] - alpha * q[row];
  const double zi = dinv[row] * ri;
  x[row] = xi;
  r[row] = ri;
  zp[row].x = zi;
  v[0] = ri * ri;
  v[1] = ri * zi;
}

// Jacobi-PCG as two launches per iteration (one lane per row: S lanes, up to
// 1024, per workgroup).  The state is kept in two
// copies: k_pcg_spmv reads st_in (written by the PCG start or the last
// update) and writes st_out, k_pcg_update the reverse, so that no launch
// reads a copy it writes.  Workgroup 0 writes the copy.  With self_sum every
// workgroup forms the scalars from the other kernel's slab partials itself
// (the same values in every workgroup); without it (matrices of many slabs,
// where nslab workgroups reading nslab partials each would cost more than a
// launch) a k_reduce launch before each kernel has already applied them to
// st_in.

// First the update's r.r / r.z (part[.][1..2]) give the iteration count, the
// convergence test and beta (nothing to sum right after the start); then
// q = A p (pcg_q_row) with partial p.q (part[.][0]).
__global__ void k_pcg_spmv(const int *__restrict__ cols, const double *__restrict__ vals,
                                                  int nrow, int nnz, int S, int nslab, double2 *__restrict__ zpA,
                                                  double2 *__restrict__ zpB, double *__restrict__ q,
                                                  double *__restrict__ part, const IterState *st_in,
                                                  IterState *st_out, int self_sum)
{
  __shared__ double lds[48];
  __shared__ IterState N;
  const int lane = threadIdx.x;
  if (lane == 0) N = *st_in;
  __syncthreads();
  if (self_sum && !N.done && !N.fresh) {
    double s[2];
    group_sum_rows<2>(part + 1, nslab, 4, s, lds);
    if (lane == 0) pcg_after_update(s[0], s[1], N.rtol, N.bnorm, N.max_its, N.its, N.rnorm, N.rz, N.beta, N.done);
    __syncthreads();
  }
  if (blockIdx.x == 0 && lane == 0) {
    IterState o = N;
    o.fresh = 0;
    *st_out = o;
  }
  if (N.done) return;
  const double beta = N.beta;
  const double2 *zc = N.pbuf ? zpB : zpA;
  double2 *zn = N.pbuf ? zpA : zpB;
  const long long row = (long long)blockIdx.x * S + lane;
  double v[1] = {0.0};
  if (row < nrow) v[0] = pcg_q_row(cols, vals, nnz, S, blockIdx.x, lane, row, zc, beta, zn, q);
  block_sum<1>(v, lds);
  if (lane == 0) part[blockIdx.x * 4] = v[0];
}

// alpha from the p.q partials; x += alpha p, r -= alpha q, z = dinv r
// (pcg_update_row); partials r.r, r.z (part[.][1..2]).
__global__ void k_pcg_update(int nrow, int S, int nslab, const double *__restrict__ dinv,
                             double2 *__restrict__ zpA, double2 *__restrict__ zpB, const double *__restrict__ q,
                             double *__restrict__ x, double *__restrict__ r, double *__restrict__ part,
                             const IterState *st_in, IterState *st_out, int self_sum)
{
  __shared__ double lds[32];
  __shared__ IterState N;
  const int lane = threadIdx.x;
  if (lane == 0) N = *st_in;
  __syncthreads();
  if (self_sum && !N.done) {
    double s[1];
    group_sum_rows<1>(part, nslab, 4, s, lds);
    if (lane == 0) {
      N.alpha = pcg_alpha(N.rz, s[0]);
      N.pbuf ^= 1;
    }
    __syncthreads();
  }
  if (blockIdx.x == 0 && lane == 0) *st_out = N;
  if (N.done) return;
  const double alpha = N.alpha;
  double2 *zp = N.pbuf ? zpB : zpA;  // the pairs whose p k_pcg_spmv wrote
  const long long row = (long long)blockIdx.x * S + lane;
  double v[2] = {0.0, 0.0};
  if (row < nrow) pcg_update_row(row, alpha, dinv, zp, q, x, r, v);
  block_sum<2>(v, lds);
  if (lane == 0) {
    part[blockIdx.x * 4 + 1] = v[0];
    part[blockIdx.x * 4 + 2] = v[1];
  }
}

// ---- the same iteration as one persistent kernel ----------------------------
// k_pcg_fused runs up to `its` PCG iterations in one cooperative launch of G
// co-resident workgroups of 256 lanes (one slab each at a time, slabs
// blockIdx.x, +G, ...).  An iteration is q = A p for the own slabs, a grid
// barrier, p.q summed by every workgroup from the slab partials, the vector
// update of the own slabs, a second barrier and the r.r / r.z sums.  Every
// workgroup sums the same partials in the same order as the two-launch
// kernels (256 lanes, slab-strided, then block_sum) and takes the same scalar
// steps, so all of them hold the same scalars, take the same branch at the
// convergence test and match k_pcg_spmv / k_pcg_update bit for bit.  Slab partials: p.q in part[.][0], r.r and r.z in part[.][1..2];
// each is rewritten only after a barrier that every reader of the previous
// value has passed.
constexpr int kFusedLanes = 256;
// Grid barrier words, each on a 256-byte line of its own (index * kBarLine):
// line 0 the release word, line 1 the count of groups arrived, line 2 + g
// the count of arrivals of group g (kBarGroup consecutive workgroups).
constexpr int kBarLine = 64;
constexpr int kBarGroup = 16;
constexpr int kBarMaxGroups = 64;
constexpr int kBarMaxGrid = kBarGroup * kBarMaxGroups;
constexpr size_t kBarBytes = (size_t)(2 + kBarMaxGroups) * kBarLine * sizeof(unsigned);
// Bound on a barrier wait in polls (each polls an agent-scope atomic, then
// sleeps): far above any legitimate wait (the slowest workgroup's phase),
// and a wait that reaches it marks the launch as faulted instead of hanging.
constexpr unsigned kBarSpinLimit = 1u << 22;

// Barrier n (1, 2, ... within a launch; the host zeroes the words before
// it).  After the workgroup barrier, lane 0 arrives at its group's counter
// with a release RMW at agent scope (the L2s of the XCDs are not coherent
// with each other, so this writes the workgroup's stores back); the last
// arrival of the group acquires and arrives at the group count; the last
// group's arrival acquires and publishes n in the release word, which every
// lane 0 polls and then acquires, before the workgroup barrier hands the
// ordering on to the other waves (HIP's own grid sync uses the same chain).
// Two levels keep the same-address atomics to 16 per group counter and one
// per group at the top, and the polls off the counters' lines.  Returns
// false when the wait gave up; the first workgroup to give up leaves the
// fault flag set (fault[0]) and the group count it saw against the count it
// waited for (fault[1], fault[2]; the IterState words after pad0).
__device__ inline bool grid_barrier(unsigned *bar, unsigned n, int *fault, int &ok_lds)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x, g = blockIdx.x / kBarGroup;
    const unsigned ngroups = (G + kBarGroup - 1) / kBarGroup;
    const unsigned gsize = (G - g * kBarGroup) < kBarGroup ? (G - g * kBarGroup) : kBarGroup;
    unsigned old = __hip_atomic_fetch_add(&bar[(2 + g) * kBarLine], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == n * gsize) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      old = __hip_atomic_fetch_add(&bar[kBarLine], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 == n * ngroups) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(&bar[0], n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    unsigned spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n) {
      if (++spins > kBarSpinLimit ||
          ((spins & 63) == 0 && __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    if (!ok) {
      __hip_atomic_store(fault + 1, (int)__hip_atomic_load(&bar[kBarLine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(fault + 2, (int)(n * ngroups), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok_lds = ok;
  }
  __syncthreads();
  return ok_lds != 0;
}

__global__ __launch_bounds__(kFusedLanes) void k_pcg_fused(
    const int *__restrict__ cols, const double *__restrict__ vals, int nrow, int nnz, int nslab,
    const double *__restrict__ dinv, double *__restrict__ x, double *__restrict__ r, double2 *__restrict__ zpA,
    double2 *__restrict__ zpB, double *__restrict__ q, double *__restrict__ part, IterState *st, unsigned *bar,
    int its)
{
  constexpr int S = kFusedLanes;
  __shared__ double lds[48];
  __shared__ double bc[2];
  __shared__ int ok_lds;
  if (st->done) return;
  const unsigned G = gridDim.x;
  const int lane = threadIdx.x;
  double beta = st->beta, rz = st->rz, alpha = st->alpha, rnorm = st->rnorm;
  const double bnorm = st->bnorm, rtol = st->rtol;
  int pbuf = st->pbuf, n = st->its, done = 0;
  const int max_its = st->max_its;
  int *fault = &st->pad0;
  unsigned arrivals = 0;
  int k = 0;
  for (; k < its && !done; k++) {
    const double2 *zc = pbuf ? zpB : zpA;
    double2 *zn = pbuf ? zpA : zpB;
    for (int slab = blockIdx.x; slab < nslab; slab += G) {
      const long long row = (long long)slab * S + lane;
      double v[1] = {0.0};
      if (row < nrow) v[0] = pcg_q_row(cols, vals, nnz, S, slab, lane, row, zc, beta, zn, q);
      block_sum<1>(v, lds);
      if (lane == 0) part[slab * 4] = v[0];
    }
    if (!grid_barrier(bar, ++arrivals, fault, ok_lds)) return;
    double s1[1], s[2];
    group_sum_rows<1>(part, nslab, 4, s1, lds);
    if (lane == 0) bc[0] = s1[0];
    __syncthreads();
    alpha = pcg_alpha(rz, bc[0]);
    pbuf ^= 1;
    double2 *zp = pbuf ? zpB : zpA;  // = zn
    for (int slab = blockIdx.x; slab < nslab; slab += G) {
      const long long row = (long long)slab * S + lane;
      double v[2] = {0.0, 0.0};
      if (row < nrow) pcg_update_row(row, alpha, dinv, zp, q, x, r, v);
      block_sum<2>(v, lds);
      if (lane == 0) {
        part[slab * 4 + 1] = v[0];
        part[slab * 4 + 2] = v[1];
      }
    }
    if (!grid_barrier(bar, ++arrivals, fault, ok_lds)) return;
    group_sum_rows<2>(part + 1, nslab, 4, s, lds);
    if (lane == 0) {
      bc[0] = s[0];
      bc[1] = s[1];
    }
    __syncthreads();
    pcg_after_update(bc[0], bc[1], rtol, bnorm, max_its, n, rnorm, rz, beta, done);
    __syncthreads();  // bc is rewritten by the next iteration
  }
  // Every workgroup read the state before its first arrival, and workgroup 0
  // writes it only after that barrier completed.
  if (blockIdx.x == 0 && lane == 0 && k > 0) {
    st->its = n;
    st->rnorm = rnorm;
    st->rz = rz;
    st->alpha = alpha;
    st->beta = beta;
    st->pbuf = pbuf;
    st->done = done;
    st->fresh = 0;
  }
}

// Zeroes the barrier words of k_pcg_fused (a kernel, not a memset node).
__global__ void k_bar_reset(unsigned *bar)
{
  if (threadIdx.x < 2 + kBarMaxGroups) bar[threadIdx.x * kBarLine] = 0u;
}

// x = 0 (a kernel, not a memset node, so that the solve can sit in a graph)
__global__ void k_zero(double *__restrict__ x, int n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// PCG work vectors of one solve
struct PcgVecs {
  double *r;
  double2 *zpA, *zpB;  // {z, p} pairs, two buffers
  double *q, *part;
  IterState *st;   // two copies (the two-launch iteration passes the state between them)
  unsigned *bar;   // grid-barrier words of k_pcg_fused
};

}  // namespace

struct ell_device_ {
  int nrow, ncol, nnz, S, nslab;
  int device;  // HIP device the mirror lives on
  int *cols;
  double *vals, *dinv;
  double *x, *b, *r, *q, *part;
  double2 *zpA, *zpB;  // PCG {z, p} pairs
  IterState *st;
  unsigned *bar;
  int fused_its;   // 0: two launches per PCG iteration; n > 0: k_pcg_fused, n iterations per launch
  int fused_grid;  // co-resident workgroups of k_pcg_fused (set with fused_its)
  hipStream_t stream;
};

static void host_diag_inv(const ell_matrix *m, std::vector<double> &dinv)
{
  dinv.resize(m->nrow);
  for (int i = 0; i < m->nrow; i++) {
    double d = 0.0;
    if (i < m->ncol) {
      const int *c = m->cols + (size_t)i * m->nnz;
      for (int j = 0; j < m->nnz; j++) {
        if (c[j] == -1) break;
        if (c[j] == i) {
          d = m->vals[(size_t)i * m->nnz + j];
          break;
        }
      }
    }
    dinv[i] = 1 / d;  // like the reference, a missing diagonal gives inf
  }
}

// The kernels read x[col] for every column id, so each must be -1 or a
// column of the matrix (ell_add_val never stores anything else).
static int check_cols(const ell_matrix *m)
{
  const size_t n = (size_t)m->nrow * m->nnz;
  for (size_t e = 0; e < n; e++) {
    const int c = m->cols[e];
    if (c < -1 || c >= m->ncol) {
      fprintf(stderr, "madmax_amd: ell column id %d at slot %zu outside [-1, %d)\n", c, e, m->ncol);
      return 1;
    }
  }
  return 0;
}

static int upload(ell_device *d, const ell_matrix *m)
{
  const size_t slots = (size_t)d->nslab * d->S * d->nnz;
  std::vector<int> hc(slots, -1);
  std::vector<double> hv(slots, 0.0);
  for (int i = 0; i < m->nrow; i++) {
    const size_t sb = (size_t)(i / d->S) * d->S * d->nnz, lane = i % d->S;
    for (int j = 0; j < m->nnz; j++) {
      hc[sb + (size_t)j * d->S + lane] = m->cols[(size_t)i * m->nnz + j];
      hv[sb + (size_t)j * d->S + lane] = m->vals[(size_t)i * m->nnz + j];
    }
  }
  std::vector<double> dinv;
  host_diag_inv(m, dinv);
  MM_HIP(hipMemcpyAsync(d->cols, hc.data(), slots * sizeof(int), hipMemcpyHostToDevice, d->stream));
  MM_HIP(hipMemcpyAsync(d->vals, hv.data(), slots * sizeof(double), hipMemcpyHostToDevice, d->stream));
  if (m->nrow)
    MM_HIP(hipMemcpyAsync(d->dinv, dinv.data(), m->nrow * sizeof(double), hipMemcpyHostToDevice,
                          d->stream));
  MM_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

extern "C" int ell_device_create(const ell_matrix *m, int slab_rows, ell_device **out)
{
  if (m == NULL || out == NULL || m->cols == NULL || m->vals == NULL) return 1;
  if (m->nrow < 0 || m->nnz < 0 || m->ncol < 0) return 2;
  int S = slab_rows > 0 ? slab_rows : 256;
  if (S % 64 != 0 || S > 1024) return 1;
  if (check_cols(m)) return 1;
  ell_device *d = (ell_device *)calloc(1, sizeof(ell_device));
  if (!d) return 2;
  d->nrow = m->nrow;
  d->ncol = m->ncol;
  d->nnz = m->nnz;
  d->S = S;
  d->nslab = (m->nrow + S - 1) / S;
  if (hipGetDevice(&d->device) != hipSuccess) {
    free(d);
    return MM_ERR_HIP;
  }
  const size_t slots = (size_t)d->nslab * S * d->nnz;
  const size_t nv = (size_t)(d->nslab ? d->nslab : 1) * S;
  const size_t nvc = (size_t)(m->ncol > (int)nv ? m->ncol : nv);
  int rc = 0;
#define ALLOC(ptr, bytes)                                                   \
  if (hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 8) != hipSuccess) { \
    rc = MM_ERR_HIP;                                                         \
    goto fail;                                                               \
  }
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    free(d);
    return MM_ERR_HIP;
  }
  ALLOC(d->cols, slots * sizeof(int));
  ALLOC(d->vals, slots * sizeof(double));
  ALLOC(d->dinv, nv * sizeof(double));
  ALLOC(d->x, nvc * sizeof(double));
  ALLOC(d->b, nv * sizeof(double));
  ALLOC(d->r, nv * sizeof(double));
  ALLOC(d->zpA, nvc * sizeof(double2));
  ALLOC(d->zpB, nvc * sizeof(double2));
  ALLOC(d->q, nv * sizeof(double));
  ALLOC(d->part, (size_t)(d->nslab + 1) * 4 * sizeof(double));
  ALLOC(d->st, 2 * sizeof(IterState));
  ALLOC(d->bar, kBarBytes);
#undef ALLOC
  if (hipMemsetAsync(d->st, 0, 2 * sizeof(IterState), d->stream) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  rc = upload(d, m);
  if (rc) goto fail;
  *out = d;
  return 0;
fail:
  ell_device_destroy(d);
  return rc;
}

extern "C" int ell_device_update(ell_device *d, const ell_matrix *m)
{
  if (!d || !m || m->nrow != d->nrow || m->nnz != d->nnz || m->ncol != d->ncol) return 1;
  if (check_cols(m)) return 1;
  return upload(d, m);
}

extern "C" int ell_device_download(const ell_device *d, int *cols, double *vals)
{
  if (!d || !cols || !vals) return 1;
  const size_t slots = (size_t)d->nslab * d->S * d->nnz;
  std::vector<int> hc(slots);
  std::vector<double> hv(slots);
  MM_HIP(hipMemcpyAsync(hc.data(), d->cols, slots * sizeof(int), hipMemcpyDeviceToHost, d->stream));
  MM_HIP(hipMemcpyAsync(hv.data(), d->vals, slots * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  MM_HIP(hipStreamSynchronize(d->stream));
  for (int i = 0; i < d->nrow; i++) {
    const size_t sb = (size_t)(i / d->S) * d->S * d->nnz, lane = i % d->S;
    for (int j = 0; j < d->nnz; j++) {
      cols[(size_t)i * d->nnz + j] = hc[sb + (size_t)j * d->S + lane];
      vals[(size_t)i * d->nnz + j] = hv[sb + (size_t)j * d->S + lane];
    }
  }
  return 0;
}

extern "C" int ell_device_time_mvp(ell_device *d, int reps, double *ms)
{
  if (!d || !ms || reps < 1) return 1;
  *ms = 0.0;
  if (d->nrow == 0) return 0;
  hipEvent_t a = nullptr, b = nullptr;
  MM_HIP(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return MM_ERR_HIP;
  }
  int rc = 0;
  k_spmv64_exact<<<d->nslab, d->S, 0, d->stream>>>(d->cols, d->vals, d->nrow, d->nnz, d->S, d->x, d->q);
  float t = 0.f;
  if (hipEventRecord(a, d->stream) != hipSuccess) rc = MM_ERR_HIP;
  for (int i = 0; i < reps && !rc; i++)
    k_spmv64_exact<<<d->nslab, d->S, 0, d->stream>>>(d->cols, d->vals, d->nrow, d->nnz, d->S, d->x, d->q);
  if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(b, d->stream) != hipSuccess ||
              hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess))
    rc = MM_ERR_HIP;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (!rc) *ms = (double)t / reps;
  return rc;
}

extern "C" int ell_device_mvp(ell_device *d, const double *x, double *y)
{
  if (!d || !x || !y) return 1;
  if (d->nrow == 0) return 0;
  MM_HIP(hipMemcpyAsync(d->x, x, d->ncol * sizeof(double), hipMemcpyHostToDevice, d->stream));
  k_spmv64_exact<<<d->nslab, d->S, 0, d->stream>>>(d->cols, d->vals, d->nrow, d->nnz, d->S, d->x, d->q);
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(y, d->q, d->nrow * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  MM_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

static int set_state(ell_device *d, IterState *st, int max_its, double rtol, double min_err)
{
  IterState h;
  memset(&h, 0, sizeof(h));
  h.max_its = max_its;
  h.rtol = rtol;
  h.min_err = min_err;
  MM_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, d->stream));
  return 0;
}

static int get_state(ell_device *d, const IterState *st, IterState *h)
{
  MM_HIP(hipMemcpyAsync(h, st, sizeof(*h), hipMemcpyDeviceToHost, d->stream));
  MM_HIP(hipStreamSynchronize(d->stream));
  return 0;
}

extern "C" int ell_device_solve_jacobi(ell_device *d, ell_solver *solver, const double *b, double *x)
{
  if (!d || !b || !x) return 1;
  if (d->nrow != d->ncol) return 1;
  if (d->nrow == 0) return 0;
  int rc = set_state(d, d->st, MAX_ITS, 0.0, MIN_ERR);
  if (rc) return rc;
  MM_HIP(hipMemcpyAsync(d->x, x, d->nrow * sizeof(double), hipMemcpyHostToDevice, d->stream));
  MM_HIP(hipMemcpyAsync(d->b, b, d->nrow * sizeof(double), hipMemcpyHostToDevice, d->stream));
  IterState h;
  for (;;) {
    for (int k = 0; k < 64; k++) {
      k_jac_res<<<d->nslab, d->S, 0, d->stream>>>(d->cols, d->vals, d->nrow, d->nnz, d->S, d->x, d->b,
                                                  d->r, d->part, d->st);
      k_reduce<<<1, 256, 0, d->stream>>>(d->part, d->nslab, RED_JAC, d->st);
      k_jac_upd<<<d->nslab, d->S, 0, d->stream>>>(d->nrow, d->S, d->dinv, d->r, d->x, d->st);
    }
    MM_HIP(hipGetLastError());
    rc = get_state(d, d->st, &h);
    if (rc) return rc;
    if (h.done) break;
  }
  MM_HIP(hipMemcpyAsync(x, d->x, d->nrow * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  MM_HIP(hipStreamSynchronize(d->stream));
  if (solver) {
    solver->its = h.its;
    solver->err = h.err;
  }
  return 0;
}

// Jacobi-PCG on the device mirror with b and x (initial guess in, solution
// out) in device memory and the work vectors V; the host polls the
// iteration state every 32 iterations.  Returns 0 when converged, 5 when
// max_its was reached.
static int pcg_run(ell_device *d, const PcgVecs &V, ell_solver *solver, const double *b, double *x)
{
  const int max_its = (solver && solver->max_its > 0) ? solver->max_its : MAX_ITS;
  const double rtol = (solver && solver->min_err > 0) ? solver->min_err : MIN_ERR;
  int rc = set_state(d, V.st, max_its, rtol, 0.0);
  if (rc) return rc;
  k_pcg_init<<<d->nslab, d->S, 0, d->stream>>>(d->cols, d->vals, d->nrow, d->nnz, d->S, x, b, d->dinv, V.r,
                                               V.zpA, V.zpB, V.part);
  k_reduce<<<1, 256, 0, d->stream>>>(V.part, d->nslab, RED_INIT, V.st);
  MM_HIP(hipGetLastError());
  IterState h;
  rc = get_state(d, V.st, &h);
  if (rc) return rc;
  if (h.bnorm == 0.0) {  // zero right-hand side: the solution is zero
    k_zero<<<(d->nrow + 255) / 256, 256, 0, d->stream>>>(x, d->nrow);
    MM_HIP(hipGetLastError());
  }
  while (!h.done && d->fused_its > 0) {
    k_bar_reset<<<1, 128, 0, d->stream>>>(V.bar);
    MM_HIP(hipGetLastError());
    const int *cols = d->cols;
    const double *vals = d->vals, *dinv = d->dinv;
    int nrow = d->nrow, nnz = d->nnz, nslab = d->nslab, its = d->fused_its;
    double *xv = x, *r = V.r, *q = V.q, *part = V.part;
    double2 *zpA = V.zpA, *zpB = V.zpB;
    IterState *st = V.st;
    unsigned *bar = V.bar;
    void *args[] = {&cols, &vals, &nrow, &nnz, &nslab, &dinv, &xv, &r, &zpA, &zpB, &q, &part, &st, &bar, &its};
    MM_HIP(hipLaunchCooperativeKernel((const void *)k_pcg_fused, dim3(d->fused_grid), dim3(kFusedLanes), args, 0,
                                      d->stream));
    rc = get_state(d, V.st, &h);
    if (rc) return rc;
    if (h.pad0) {
      fprintf(stderr,
              "madmax_amd: k_pcg_fused: a grid barrier wait gave up at %d of %d group arrivals (%d workgroups)\n",
              h.pad1, h.pad2, d->fused_grid);
      return MM_ERR_HIP;
    }
  }
  // the state passes st[0] -> st[1] -> st[0]; up to kSelfSumSlabs slabs two
  // launches per iteration, beyond that a one-workgroup k_reduce before each
  const int self_sum = d->nslab <= kSelfSumSlabs ? 1 : 0;
  while (!h.done) {
    for (int k = 0; k < 32; k++) {
      k_pcg_spmv<<<d->nslab, d->S, 0, d->stream>>>(d->cols, d->vals, d->nrow, d->nnz, d->S, d->nslab, V.zpA,
                                                   V.zpB, V.q, V.part, V.st, V.st + 1, self_sum);
      if (!self_sum) k_reduce<<<1, 256, 0, d->stream>>>(V.part, d->nslab, RED_PQ, V.st + 1);
      k_pcg_update<<<d->nslab, d->S, 0, d->stream>>>(d->nrow, d->S, d->nslab, d->dinv, V.zpA, V.zpB, V.q, x,
                                                     V.r, V.part, V.st + 1, V.st, self_sum);
      if (!self_sum) k_reduce<<<1, 256, 0, d->stream>>>(V.part, d->nslab, RED_RR, V.st);
    }
    MM_HIP(hipGetLastError());
    rc = get_state(d, V.st, &h);
    if (rc) return rc;
  }
  MM_HIP(hipStreamSynchronize(d->stream));
  const double relres = (h.bnorm == 0.0) ? 0.0 : h.rnorm / h.bnorm;
  if (solver) {
    solver->its = h.its;
    solver->err = relres;
  }
  return (h.bnorm == 0.0 || h.rnorm <= rtol * h.bnorm) ? 0 : MM_NOT_CONVERGED;
}

extern "C" int ell_device_set_pcg_fused(ell_device *d, int its_per_launch)
{
  if (!d || its_per_launch < 0) return 1;
  if (its_per_launch == 0) {
    d->fused_its = 0;
    return 0;
  }
  if (d->S != kFusedLanes || d->nslab == 0) return 1;
  int coop = 0, ncu = 0, per_cu = 0;
  MM_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, d->device));
  if (!coop) return 1;
  MM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d->device));
  MM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pcg_fused, kFusedLanes, 0));
  if (ncu < 1 || per_cu < 1) return 1;
  // Half of what fits: a grid at the full occupancy limit needs every wave
  // slot of every CU free at once, and the launch admits it even when some
  // are not (its barrier waits then run out).  Workgroups loop over slabs,
  // so a smaller grid only lengthens each one's share.
  long long cap = (long long)ncu * (per_cu > 1 ? per_cu / 2 : 1);
  if (cap > kBarMaxGrid) cap = kBarMaxGrid;
  d->fused_grid = (int)(d->nslab < cap ? d->nslab : cap);
  d->fused_its = its_per_launch;
  return 0;
}

extern "C" int ell_device_solve_pcg(ell_device *d, ell_solver *solver, const double *b, double *x)
{
  if (!d || !b || !x) return 1;
  if (d->nrow != d->ncol) return 1;
  if (d->nrow == 0) return 0;
  MM_HIP(hipMemcpyAsync(d->x, x, d->nrow * sizeof(double), hipMemcpyHostToDevice, d->stream));
  MM_HIP(hipMemcpyAsync(d->b, b, d->nrow * sizeof(double), hipMemcpyHostToDevice, d->stream));
  const PcgVecs V{d->r, d->zpA, d->zpB, d->q, d->part, d->st, d->bar};
  const int rc = pcg_run(d, V, solver, d->b, d->x);
  if (rc != 0 && rc != MM_NOT_CONVERGED) return rc;
  MM_HIP(hipMemcpyAsync(x, d->x, d->nrow * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  MM_HIP(hipStreamSynchronize(d->stream));
  return rc;
}

// ---- internal: a device ELL filled by kernels (fe.hip) --------------------

int mm::ell_device_alloc(int nrow, int nnz, int S, ell_device **out)
{
  if (!out || nrow < 0 || nnz < 1 || S % 64 != 0 || S < 64 || S > 1024) return MM_ERR_ARG;
  ell_device *d = (ell_device *)calloc(1, sizeof(ell_device));
  if (!d) return MM_ERR_ALLOC;
  d->nrow = d->ncol = nrow;
  d->nnz = nnz;
  d->S = S;
  d->nslab = (nrow + S - 1) / S;
  if (hipGetDevice(&d->device) != hipSuccess) {
    free(d);
    return MM_ERR_HIP;
  }
  const size_t slots = (size_t)d->nslab * S * nnz;
  const size_t nv = (size_t)(d->nslab ? d->nslab : 1) * S;
  int rc = 0;
#define ALLOC(ptr, bytes)                                                   \
  if (hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 8) != hipSuccess) { \
    rc = MM_ERR_HIP;                                                         \
    goto fail;                                                               \
  }
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    free(d);
    return MM_ERR_HIP;
  }
  ALLOC(d->cols, slots * sizeof(int));
  ALLOC(d->vals, slots * sizeof(double));
  ALLOC(d->dinv, nv * sizeof(double));
  ALLOC(d->x, nv * sizeof(double));
  ALLOC(d->b, nv * sizeof(double));
  ALLOC(d->r, nv * sizeof(double));
  ALLOC(d->zpA, nv * sizeof(double2));
  ALLOC(d->zpB, nv * sizeof(double2));
  ALLOC(d->q, nv * sizeof(double));
  ALLOC(d->part, (size_t)(d->nslab + 1) * 4 * sizeof(double));
  ALLOC(d->st, 2 * sizeof(IterState));
  ALLOC(d->bar, kBarBytes);
#undef ALLOC
  if (hipMemsetAsync(d->cols, 0xff, slots * sizeof(int), d->stream) != hipSuccess ||
      hipMemsetAsync(d->vals, 0, slots * sizeof(double), d->stream) != hipSuccess ||
      hipMemsetAsync(d->st, 0, 2 * sizeof(IterState), d->stream) != hipSuccess ||
      hipStreamSynchronize(d->stream) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  *out = d;
  return 0;
fail:
  ell_device_destroy(d);
  return rc;
}

void mm::ell_device_arrays(ell_device *d, int **cols, double **vals, double **dinv, int *S, hipStream_t *s)
{
  *cols = d->cols;
  *vals = d->vals;
  *dinv = d->dinv;
  *S = d->S;
  *s = d->stream;
}

int mm::ell_device_pcg_device(ell_device *d, ell_solver *solver, const double *b, double *x)
{
  if (!d || !b || !x) return MM_ERR_ARG;
  if (d->nrow == 0) return 0;
  const PcgVecs V{d->r, d->zpA, d->zpB, d->q, d->part, d->st, d->bar};
  return pcg_run(d, V, solver, b, x);
}

// ---- the same solve without any allocation ----------------------------------
// The PCG work vectors, partial sums and iteration state of
// ell_device_solve_pcg_noalloc live in this file-scope array: six vectors of
// ELL_NOALLOC_MAX_ROWS doubles, the partials of up to ELL_NOALLOC_MAX_ROWS / 64
// slabs, two IterState copies and the fused kernel's barrier words (about
// 48.5 MiB, held on every GPU the library runs on until the process ends).
// Host code never uses the array's name as an address: the per-device
// address comes from hipGetSymbolAddress.
namespace {
constexpr size_t kNaRows = ELL_NOALLOC_MAX_ROWS;
constexpr size_t kNaPart = (kNaRows / 64 + 1) * 4;
constexpr size_t kNaState = (2 * sizeof(IterState) + 7) / 8;
constexpr size_t kNaBar = kBarBytes / 8;  // the grid-barrier words of k_pcg_fused
constexpr size_t kNaDoubles = 6 * kNaRows + kNaPart + kNaState + kNaBar;
constexpr int kNaMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNaDoubles];

static std::mutex g_na_mu;
static double *g_na_ptr[kNaMaxDevices];

// Address of `scratch` on `device`, looked up once per device with that
// device current.
static int noalloc_scratch(int device, double **out)
{
  if (device < 0 || device >= kNaMaxDevices) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[device]) {
    int cur = -1;
    MM_HIP(hipGetDevice(&cur));
    if (cur != device) MM_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) MM_HIP(hipSetDevice(cur));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "madmax_amd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorName(e));
      return MM_ERR_HIP;
    }
    g_na_ptr[device] = (double *)p;
  }
  *out = g_na_ptr[device];
  return 0;
}

extern "C" int ell_device_solve_pcg_noalloc(ell_device *d, ell_solver *solver, const double *b, double *x)
{
  if (!d || !b || !x) return 1;
  if (d->nrow != d->ncol) return 1;
  if ((size_t)d->nrow > kNaRows) {
    fprintf(stderr, "madmax_amd: ell_device_solve_pcg_noalloc: %d rows exceed the %zu-row scratch\n", d->nrow,
            kNaRows);
    return 1;
  }
  if (d->nrow == 0) return 0;
  double *sp = nullptr;
  const int rc = noalloc_scratch(d->device, &sp);
  if (rc) return rc;
  // r, then the two pair buffers (2 rows' worth each, 16-byte aligned), q, the
  // partials, the two states and the barrier words
  const PcgVecs V{sp, reinterpret_cast<double2 *>(sp + kNaRows), reinterpret_cast<double2 *>(sp + 3 * kNaRows),
                  sp + 5 * kNaRows, sp + 6 * kNaRows, reinterpret_cast<IterState *>(sp + 6 * kNaRows + kNaPart),
                  reinterpret_cast<unsigned *>(sp + 6 * kNaRows + kNaPart + kNaState)};
  return pcg_run(d, V, solver, b, x);
}

extern "C" int ell_device_destroy(ell_device *d)
{
  if (!d) return 1;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  (void)hipFree(d->cols);
  (void)hipFree(d->vals);
  (void)hipFree(d->dinv);
  (void)hipFree(d->x);
  (void)hipFree(d->b);
  (void)hipFree(d->r);
  (void)hipFree(d->zpA);
  (void)hipFree(d->zpB);
  (void)hipFree(d->q);
  (void)hipFree(d->part);
  (void)hipFree(d->st);
  (void)hipFree(d->bar);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  free(d);
  return 0;
}

// ---- reference entry points running on the GPU ------------------------------

// src/ell.c:127-141
extern "C" int ell_mvp(ell_matrix *m, double *x, double *y)
{
  if (m == NULL || x == NULL || y == NULL) return 1;
  ell_device *d = NULL;
  int rc = ell_device_create(m, 0, &d);
  if (rc) return rc;
  rc = ell_device_mvp(d, x, y);
  ell_device_destroy(d);
  return rc;
}

// src/ell.c:143-184
extern "C" int ell_solve_jacobi(ell_solver *solver, ell_matrix *m, double *b, double *x)
{
  if (m == NULL || b == NULL || x == NULL) return 1;
  ell_device *d = NULL;
  int rc = ell_device_create(m, 0, &d);
  if (rc) return rc;
  rc = ell_device_solve_jacobi(d, solver, b, x);
  ell_device_destroy(d);
  return rc;
}

extern "C" int ell_solve_pcg(ell_solver *solver, ell_matrix *m, double *b, double *x)
{
  if (m == NULL || b == NULL || x == NULL) return 1;
  ell_device *d = NULL;
  int rc = ell_device_create(m, 0, &d);
  if (rc) return rc;
  rc = ell_device_solve_pcg(d, solver, b, x);
  ell_device_destroy(d);
  return rc;
}
