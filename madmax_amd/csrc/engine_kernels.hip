// engine_kernels.hip -- CDNA4 kernels of the RVE engine.
//
// Arena layout in HBM.  Rows (dofs) are numbered 3*node + d; a slab holds SN
// consecutive nodes (3*SN rows) of one RVE and is processed by one workgroup
// of SN lanes, one lane per node:
//   vals  : [slab][81][3*SN]  ELL values in the reference's slot order
//                             (slot j = 3t + d' of row 3n + d), column-major
//                             inside the slab so lane i reads the three rows
//                             of its node as 12 contiguous bytes
//   bcols : [slab][27][SN]    RVE-local node id of slot block t (-1 = empty);
//                             the reference's column of slot 3t + d' is
//                             3*bcols + d', shared by the node's three rows
//   vectors [row][KS]         x64 (FP64 displacement), d (correction), r, z,
//                             p, q (PCG), KS = load cases padded to even
//   dinv  : [row]             Jacobi preconditioner
//   part  : [slab][96]        per-workgroup partial sums (FP64)
// Every resident RVE owns a contiguous run of slabs; its first row is
// first_slab * 3 * SN.
//
// The solver kernels advance every resident RVE by one phase step; their grid
// covers all slabs and each workgroup reads the phase of its RVE.  The
// workgroup that finishes last for an RVE (arrive_last) reduces that RVE's
// partials in slab order, so an RVE's results do not depend on which other
// RVEs share the launch.
#include <hip/hip_runtime.h>

#include "mm_device.h"
#include "mm_internal.h"

namespace mm {

__constant__ SlotTables g_slots;

int upload_slot_tables(hipStream_t s)
{
  SlotTables t;
  build_slot_tables(&t);
  MM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_slots), &t, sizeof(t), 0, hipMemcpyHostToDevice, s));
  MM_HIP(hipStreamSynchronize(s));
  return 0;
}

constexpr int kPart = kPartDoubles;

__host__ __device__ constexpr int vec_stride(int K) { return (K == 1) ? 1 : ((K + 1) & ~1); }

__device__ __forceinline__ int node_class(int ix, int iy, int iz, int nx, int ny, int nz)
{
  const int cx = (ix == 0) ? 0 : ((ix == nx - 1) ? 2 : 1);
  const int cy = (iy == 0) ? 0 : ((iy == ny - 1) ? 2 : 1);
  const int cz = (iz == 0) ? 0 : ((iz == nz - 1) ? 2 : 1);
  return cx + 3 * cy + 9 * cz;
}

// ---------------------------------------------------------------------------
// Admission: scatter new RVE states and their slab ownership.
__global__ void k_admit(int n, const int *__restrict__ slots, const RveDev *__restrict__ states,
                        RveDev *__restrict__ rves, int *__restrict__ slab_rve,
                        int *__restrict__ slab_node0, int SN)
{
  const int i = blockIdx.x;
  if (i >= n) return;
  const int slot = slots[i];
  const RveDev &st = states[i];
  if (threadIdx.x == 0) rves[slot] = st;
  for (int s = threadIdx.x; s < st.nslab; s += blockDim.x) {
    slab_rve[st.first_slab + s] = slot;
    slab_node0[st.first_slab + s] = s * SN;
  }
}

// Element stiffness per (RVE, material), one wavefront each.  B at the 8
// gauss points is built in LDS from the hex8 tables of src/fem.c:309-405
// with dN/dx_d = ds * 2 / h_d; C is the isotropic tangent in the form of
// src/material.c:16-22; Ke follows micro/src/assembly.c:97-107 term by term
// (terms with an exactly zero factor are skipped, which leaves every sum
// bit-identical).  Also stores C and sum_gp B w for the averaging phase.
__global__ __launch_bounds__(64) void k_ke(int n, const int *__restrict__ work,
                                           const RveDev *__restrict__ rves,
                                           double *__restrict__ ke_tab, double *__restrict__ c_tab_out,
                                           double *__restrict__ bw_tab)
{
#pragma clang fp contract(off)
  __shared__ double B[6][24][8];
  __shared__ double C[36];
  __shared__ double wp[8];
  if ((int)blockIdx.x >= n) return;
  const int slot = work[2 * blockIdx.x], mat = work[2 * blockIdx.x + 1];
  const RveDev &R = rves[slot];
  const int tid = threadIdx.x;
  const double h[3] = {R.hx, R.hy, R.hz};
  for (int i = tid; i < 6 * 24 * 8; i += 64) (&B[0][0][0])[i] = 0.0;
  __syncthreads();
  const double ga = 0.577350269189626;
  // sign pattern of nodes / gauss points: 0(-,-,-) 1(+,-,-) 2(+,+,-) 3(-,+,-) 4..7 with z = +
  for (int i = tid; i < 8 * 8; i += 64) {
    const int a = i >> 3, g = i & 7;
    const double sa[3] = {(a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0,
                          (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0, (a >= 4) ? 1.0 : -1.0};
    const double sg[3] = {(g == 1 || g == 2 || g == 5 || g == 6) ? 1.0 : -1.0,
                          (g == 2 || g == 3 || g == 6 || g == 7) ? 1.0 : -1.0, (g >= 4) ? 1.0 : -1.0};
    double f[3];
    for (int d = 0; d < 3; d++) f[d] = 1.0 + sa[d] * (sg[d] * ga);
    const double ds0 = sa[0] * f[1] * f[2] * 0.125;
    const double ds1 = sa[1] * f[0] * f[2] * 0.125;
    const double ds2 = sa[2] * f[0] * f[1] * 0.125;
    const double dx = ds0 * 2 / h[0], dy = ds1 * 2 / h[1], dz = ds2 * 2 / h[2];
    B[0][3 * a + 0][g] = dx;
    B[1][3 * a + 1][g] = dy;
    B[2][3 * a + 2][g] = dz;
    B[3][3 * a + 0][g] = dy;
    B[3][3 * a + 1][g] = dx;
    B[4][3 * a + 1][g] = dz;
    B[4][3 * a + 2][g] = dy;
    B[5][3 * a + 0][g] = dz;
    B[5][3 * a + 2][g] = dx;
  }
  if (tid < 36) {
    const int i = tid / 6, j = tid % 6;
    const double nu = R.poisson[mat], E = R.young[mat];
    double base = 0.0;
    if (i < 3 && j < 3) base = (i == j) ? 1.0 - nu : nu;
    else if (i == j) base = (1 - 2 * nu) / 2;
    C[tid] = base * (E / ((1 + nu) * (1 - 2 * nu)));
  }
  if (tid < 8) {
    const double vol = h[0] * h[1] * h[2];
    wp[tid] = vol / 8;
  }
  __syncthreads();
  double *ke = ke_tab + ((size_t)slot * kNmat + mat) * kKe;
  for (int e = tid; e < kKe; e += 64) {
    const int i = e / 24, j = e % 24;
    double acc = 0.0;
    for (int g = 0; g < 8; g++)
      for (int k = 0; k < 6; k++)
        for (int hh = 0; hh < 6; hh++) {
          const double bh = B[hh][i][g];
          const double c = C[hh * 6 + k];
          const double bk = B[k][j][g];
          if (bh == 0.0 || c == 0.0 || bk == 0.0) continue;
          acc = acc + bh * c * bk * wp[g];
        }
    ke[e] = acc;
  }
  if (tid < 36) c_tab_out[((size_t)slot * kNmat + mat) * 36 + tid] = C[tid];
  if (mat == 0)
    for (int e = tid; e < 6 * 24; e += 64) {
      const int v = e / 24, i = e % 24;
      double s = 0.0;
      for (int g = 0; g < 8; g++) s += B[v][i][g] * wp[g];
      bw_tab[(size_t)slot * 144 + e] = s;
    }
}

// Phase map: an element is fiber when the in-plane distance of its centroid
// to a fiber axis is <= r (micro/src/micro_struct.c:114-127).
__global__ void k_phase(int nblk, const int2 *__restrict__ work, const RveDev *__restrict__ rves,
                        uint8_t *__restrict__ elem_mat)
{
#pragma clang fp contract(off)
  if ((int)blockIdx.x >= nblk) return;
  const int2 w = work[blockIdx.x];
  const RveDev &R = rves[w.x];
  const int e = w.y + threadIdx.x;
  if (e >= R.nelm) return;
  const int ex = e % R.nex, ey = (e / R.nex) % R.ney;
  const double cx = (ex + 0.5) * R.hx, cy = (ey + 0.5) * R.hy;
  uint8_t m = 0;
  for (int f = 0; f < R.nfib; f++) {
    const double dx = cx - R.fx[f], dy = cy - R.fy[f];
    double l = 0.0;
    l = l + dx * dx;
    l = l + dy * dy;
    l = sqrt(l);
    if (l <= R.fr[f]) {
      m = 1;
      break;
    }
  }
  elem_mat[R.elem_base + e] = m;
}

// ELL assembly: one lane per node, writing the three rows of the node.  Every
// slot value is the sum of the Ke entries of the elements sharing the two
// nodes, in ascending element id, the first stored and the rest added
// (src/ell.c:46-73, micro/src/assembly.c:93-110).  Boundary rows become
// identity rows with their columns kept (src/ell.c:86-118); the empty tail
// holds -1 / 0.  Also writes the Jacobi diagonal inverse and the initial
// displacement: affine on the boundary (micro/src/homogenize.c:227-242,
// 335-342), zero or the kept values inside.
// mode 0: full assembly; mode 1: displacement only (resident matrix reused).
template <typename TV, int K>
__global__ void k_assemble(int nblk, const int *__restrict__ slab_list, const int *__restrict__ slab_rve,
                           const int *__restrict__ slab_node0, const RveDev *__restrict__ rves,
                           const double *__restrict__ ke_tab, const uint8_t *__restrict__ elem_mat,
                           int *__restrict__ bcols, TV *__restrict__ vals, TV *__restrict__ dinv,
                           double *__restrict__ x64, int SN, int mode)
{
#pragma clang fp contract(off)
  constexpr int KS = vec_stride(K);
  if ((int)blockIdx.x >= nblk) return;
  const int slab = slab_list[blockIdx.x];
  const int slot = slab_rve[slab];
  const RveDev &R = rves[slot];
  const int tid = threadIdx.x;
  const int n = slab_node0[slab] + tid;
  const int RS = 3 * SN;
  const long long vb = (long long)slab * kW * RS;
  const long long cb = (long long)slab * 27 * SN;
  const long long grow0 = R.row_base + 3LL * n;
  const int nn = R.nrow / 3;
  if (n >= nn) {
    if (mode == 0) {
      for (int t = 0; t < 27; t++) bcols[cb + (long long)t * SN + tid] = -1;
      for (int j = 0; j < kW; j++)
        for (int d = 0; d < 3; d++) vals[vb + (long long)j * RS + 3 * tid + d] = (TV)0;
      for (int d = 0; d < 3; d++) dinv[grow0 + d] = (TV)0;
    }
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < KS; k++) x64[(grow0 + d) * KS + k] = 0.0;
    return;
  }
  const int nx = R.nx, nxy = R.nx * R.ny;
  const int ix = n % nx, iy = (n / nx) % R.ny, iz = n / nxy;
  const int cls = node_class(ix, iy, iz, nx, R.ny, R.nz);
  const bool bc = (cls != 13);
  if (mode == 0) {
    const int ns = g_slots.nslot[cls];
    const int dslot = g_slots.diag_slot[cls];
    const double *ke0 = ke_tab + (size_t)slot * kNmat * kKe;
    double diag[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < 27; t++) {
      if (t < ns) {
        const int key = g_slots.off[cls][t];
        const int m = n + (key % 3 - 1) + ((key / 3) % 3 - 1) * nx + (key / 9 - 1) * nxy;
        double v[3][3];  // [d][d']
        if (bc) {
          for (int d = 0; d < 3; d++)
            for (int dd = 0; dd < 3; dd++) v[d][dd] = (m == n && d == dd) ? 1.0 : 0.0;
        } else {
          const int nc = g_slots.ncontrib[cls][t];
          for (int c = 0; c < nc; c++) {
            const int code = g_slots.contrib[cls][t][c];
            const int ox = (code & 1) - 1, oy = ((code >> 1) & 1) - 1, oz = ((code >> 2) & 1) - 1;
            const int a = (code >> 3) & 7, b = (code >> 6) & 7;
            const int e = (ix + ox) + (iy + oy) * R.nex + (iz + oz) * R.nex * R.ney;
            const double *ke = ke0 + (size_t)elem_mat[R.elem_base + e] * kKe + 3 * a * kNd + 3 * b;
            for (int d = 0; d < 3; d++)
              for (int dd = 0; dd < 3; dd++)
                v[d][dd] = (c == 0) ? ke[d * kNd + dd] : v[d][dd] + ke[d * kNd + dd];
          }
        }
        if (t == dslot)
          for (int d = 0; d < 3; d++) diag[d] = v[d][d];
        bcols[cb + (long long)t * SN + tid] = m;
        for (int dd = 0; dd < 3; dd++)
          for (int d = 0; d < 3; d++) vals[vb + (long long)(3 * t + dd) * RS + 3 * tid + d] = (TV)v[d][dd];
      } else {
        bcols[cb + (long long)t * SN + tid] = -1;
        for (int dd = 0; dd < 3; dd++)
          for (int d = 0; d < 3; d++) vals[vb + (long long)(3 * t + dd) * RS + 3 * tid + d] = (TV)0;
      }
    }
    for (int d = 0; d < 3; d++) dinv[grow0 + d] = (TV)(1.0 / diag[d]);
  }
  // initial displacement for every load case
  const double X[3] = {ix * R.hx, iy * R.hy, iz * R.hz};
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < KS; k++) {
      if (k < R.nload && bc) {
        const double *s = R.load[k];
        double u;
        if (d == 0) u = s[0] * X[0] + s[3] / 2 * X[1] + s[5] / 2 * X[2];
        else if (d == 1) u = s[3] / 2 * X[0] + s[1] * X[1] + s[4] / 2 * X[2];
        else u = s[5] / 2 * X[0] + s[4] / 2 * X[1] + s[2] * X[2];
        x64[(grow0 + d) * KS + k] = u;
      } else if (!(R.warm && k < R.nload)) {
        x64[(grow0 + d) * KS + k] = 0.0;
      }
    }
}

template <typename T>
struct Vec2Of;
template <>
struct Vec2Of<float> { using type = float2; };
template <>
struct Vec2Of<double> { using type = double2; };

// Loads / stores N consecutive values (N even -> two-wide vector accesses).
template <typename TS, int N>
__device__ __forceinline__ void load_vec(const TS *__restrict__ p, TS (&v)[N])
{
  if constexpr (N % 2 == 0) {
    using V2 = typename Vec2Of<TS>::type;
    const V2 *q = reinterpret_cast<const V2 *>(p);
#pragma unroll
    for (int k = 0; k < N / 2; k++) {
      const V2 t = q[k];
      v[2 * k] = t.x;
      v[2 * k + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = p[k];
  }
}
template <typename TS, int N>
__device__ __forceinline__ void store_vec(TS *__restrict__ p, const TS (&v)[N])
{
  if constexpr (N % 2 == 0) {
    using V2 = typename Vec2Of<TS>::type;
    V2 *q = reinterpret_cast<V2 *>(p);
#pragma unroll
    for (int k = 0; k < N / 2; k++) q[k] = V2{v[2 * k], v[2 * k + 1]};
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = v[k];
  }
}

// w = A z for the three rows of one node and KS load cases: 27 slot blocks,
// each one column id, a 3x3 value block and one gathered 3*KS vector run.
template <typename TS, int KS>
__device__ __forceinline__ void node_spmv(const int *__restrict__ bcols, const TS *__restrict__ vals,
                                          long long cb, long long vb, int SN, int tid, long long rbase,
                                          const TS *__restrict__ z, TS (&w)[3][KS])
{
  const int RS = 3 * SN;
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int k = 0; k < KS; k++) w[d][k] = (TS)0;
#pragma unroll 3
  for (int t = 0; t < 27; t++) {
    const int m = __builtin_nontemporal_load(&bcols[cb + (long long)t * SN + tid]);
    TS v[3][3];  // [d'][d]
#pragma unroll
    for (int dd = 0; dd < 3; dd++)
#pragma unroll
      for (int d = 0; d < 3; d++)
        v[dd][d] = __builtin_nontemporal_load(&vals[vb + (long long)(3 * t + dd) * RS + 3 * tid + d]);
    if (m >= 0) {
      TS xz[3 * KS];
      load_vec<TS, 3 * KS>(z + (rbase + 3LL * m) * KS, xz);
#pragma unroll
      for (int dd = 0; dd < 3; dd++)
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int k = 0; k < KS; k++) w[d][k] = fma(v[dd][d], xz[dd * KS + k], w[d][k]);
    }
  }
}

// ---------------------------------------------------------------------------
// Kernel A: PH_INNER and PH_XUPD.
// INNER: w = A z, then p = z + beta p and q = w + beta q (so q = A p without
// gathering p), and p.q; the last workgroup forms alpha = r.z / p.q.
// XUPD: x64 += d.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_a(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[K * 16];
  const int slab = blockIdx.x, tid = threadIdx.x, SN = P.SN;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  const int phase = R->phase;
  if (phase != PH_INNER && phase != PH_XUPD) return;
  const int n = P.slab_node0[slab] + tid;
  const bool valid = n < R->nrow / 3;
  const long long rbase = R->row_base;
  const long long grow0 = rbase + 3LL * n;
  const int nslab = R->nslab;
  double *part = P.part + (size_t)slab * kPart;

  if (phase == PH_INNER) {
    TS *p = (TS *)P.p;
    TS *q = (TS *)P.q;
    TS beta[KS];
#pragma unroll
    for (int k = 0; k < KS; k++) beta[k] = (k < K) ? (TS)R->beta[k] : (TS)0;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = 0.0;
    if (valid) {
      TS w[3][KS];
      node_spmv<TS, KS>(P.bcols, (const TS *)P.vals, (long long)slab * 27 * SN, (long long)slab * kW * 3 * SN,
                        SN, tid, rbase, (const TS *)P.z, w);
      TS zi[3 * KS], pi[3 * KS], qi[3 * KS];
      load_vec<TS, 3 * KS>((const TS *)P.z + grow0 * KS, zi);
      load_vec<TS, 3 * KS>(p + grow0 * KS, pi);
      load_vec<TS, 3 * KS>(q + grow0 * KS, qi);
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int k = 0; k < KS; k++) {
          const int i = d * KS + k;
          pi[i] = fma(beta[k], pi[i], zi[i]);
          qi[i] = fma(beta[k], qi[i], w[d][k]);
          if (k < K) v[k < K ? k : 0] += (double)pi[i] * (double)qi[i];
        }
      store_vec<TS, 3 * KS>(p + grow0 * KS, pi);
      store_vec<TS, 3 * KS>(q + grow0 * KS, qi);
    }
    block_sum<K>(v, lds);
    if (tid == 0)
      for (int k = 0; k < K; k++) part[k] = v[k];
    if (arrive_last(&R->cntA, nslab)) {
      double s[K];
      group_reduce<K>(P.part + (size_t)R->first_slab * kPart, nslab, kPart, 0, s, lds);
      if (tid == 0)
        for (int k = 0; k < K; k++) {
          const bool done = (R->done_in >> k) & 1u;
          R->alpha[k] = (done || s[k] == 0.0) ? 0.0 : R->rz[k] / s[k];
        }
    }
    return;
  }

  // PH_XUPD
  if (valid) {
    const TS *d = (const TS *)P.d;
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int k = 0; k < K; k++) P.x64[(grow0 + r) * KS + k] += (double)d[(grow0 + r) * KS + k];
  }
  if (arrive_last(&R->cntA, nslab) && tid == 0) {
    R->sweep += 1;
    R->phase = PH_RES;
  }
}

// ---------------------------------------------------------------------------
// Kernel B: PH_INNER (d += alpha p, r -= alpha q, z = dinv r, r.r, r.z); the
// last workgroup checks the per-load-case convergence and forms beta.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_b(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[2 * K * 16];
  const int slab = blockIdx.x, tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  if (R->phase != PH_INNER) return;
  const int n = P.slab_node0[slab] + tid;
  const bool valid = n < R->nrow / 3;
  const long long grow0 = R->row_base + 3LL * n;
  const int nslab = R->nslab;
  double *part = P.part + (size_t)slab * kPart;
  double v[2 * K];
#pragma unroll
  for (int i = 0; i < 2 * K; i++) v[i] = 0.0;
  if (valid) {
    TS *dv = (TS *)P.d, *rv = (TS *)P.r, *zv = (TS *)P.z;
    const TS *p = (const TS *)P.p, *q = (const TS *)P.q, *dinv = (const TS *)P.dinv;
    TS alpha[KS];
#pragma unroll
    for (int k = 0; k < KS; k++) alpha[k] = (k < K) ? (TS)R->alpha[k] : (TS)0;
    TS di[3 * KS], ri[3 * KS], pi[3 * KS], qi[3 * KS], zi[3 * KS];
    load_vec<TS, 3 * KS>(dv + grow0 * KS, di);
    load_vec<TS, 3 * KS>(rv + grow0 * KS, ri);
    load_vec<TS, 3 * KS>(p + grow0 * KS, pi);
    load_vec<TS, 3 * KS>(q + grow0 * KS, qi);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const TS dg = dinv[grow0 + d];
#pragma unroll
      for (int k = 0; k < KS; k++) {
        const int i = d * KS + k;
        di[i] = fma(alpha[k], pi[i], di[i]);
        ri[i] = fma(-alpha[k], qi[i], ri[i]);
        zi[i] = dg * ri[i];
        if (k < K) {
          v[k < K ? k : 0] += (double)ri[i] * (double)ri[i];
          v[K + (k < K ? k : 0)] += (double)ri[i] * (double)zi[i];
        }
      }
    }
    store_vec<TS, 3 * KS>(dv + grow0 * KS, di);
    store_vec<TS, 3 * KS>(rv + grow0 * KS, ri);
    store_vec<TS, 3 * KS>(zv + grow0 * KS, zi);
  }
  block_sum<2 * K>(v, lds);
  if (tid == 0)
    for (int i = 0; i < 2 * K; i++) part[K + i] = v[i];
  if (!arrive_last(&R->cntB, nslab)) return;
  double s2[2 * K];
  group_reduce<2 * K>(P.part + (size_t)R->first_slab * kPart, nslab, kPart, K, s2, lds);
  if (tid == 0) {
    // each load case stops at its own sweep target; the sweep ends when all have
    R->its += 1;
    unsigned done = R->done_in;
    const unsigned all = (1u << K) - 1u;
    for (int k = 0; k < K; k++) {
      const double rr = s2[k], rz = s2[K + k];
      if ((done >> k) & 1u) {
        R->beta[k] = 0.0;
        continue;
      }
      if (sqrt(rr) <= R->bnorm_in[k] || R->rz[k] == 0.0) {
        done |= 1u << k;
        R->beta[k] = 0.0;
      } else {
        R->beta[k] = rz / R->rz[k];
        R->rz[k] = rz;
      }
    }
    R->done_in = done;
    if ((done & all) == all || R->its >= sp.max_inner_its) {
      R->its_total += R->its;
      R->phase = PH_XUPD;
    }
  }
}

// ---------------------------------------------------------------------------
// Residual kernel: PH_RES, the FP64 residual r = -K x64 computed matrix-free
// from the element stiffness tables (the Newton / refinement residual of
// micro/src/assembly.c:33-82; boundary rows keep a zero residual, :59-74),
// then the start of the next PCG sweep (d = 0, z = dinv r, p = q = 0): every
// refinement sweep solves the correction equation from zero.
// The relative target refers to the cold-start right-hand side (interior
// displacement zero, the ||b|| of the first solve); a warm-started RVE
// measures it at sweep 0 from the boundary neighbours alone.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_res(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[4 * K * 16];
  const int slab = blockIdx.x, tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  if (R->phase != PH_RES) return;
  const int n = P.slab_node0[slab] + tid;
  const bool valid = n < R->nrow / 3;
  const long long rbase = R->row_base;
  const int nslab = R->nslab;
  double *part = P.part + (size_t)slab * kPart;
  const bool need_cold = R->warm && R->sweep == 0;
  double v[4 * K];
#pragma unroll
  for (int i = 0; i < 4 * K; i++) v[i] = 0.0;
  if (valid) {
    TS *dv = (TS *)P.d, *rv = (TS *)P.r, *zv = (TS *)P.z;
    const TS *dinv = (const TS *)P.dinv;
    const int nx = R->nx, nxy = R->nx * R->ny;
    const int ix = n % nx, iy = (n / nx) % R->ny, iz = n / nxy;
    const int cls = node_class(ix, iy, iz, nx, R->ny, R->nz);
    const double *ke0 = P.ke_tab + (size_t)slot * kNmat * kKe;
#pragma unroll 1
    for (int d = 0; d < 3; d++) {
      const long long grow = rbase + 3LL * n + d;
      double r64[K], rc64[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        r64[k] = 0.0;
        rc64[k] = 0.0;
      }
      if (cls == 13) {
        const int ne = g_slots.nelem[cls];
#pragma unroll 1
        for (int i = 0; i < ne; i++) {
          const int code = g_slots.elem[cls][i];
          const int ox = (code & 1) - 1, oy = ((code >> 1) & 1) - 1, oz = ((code >> 2) & 1) - 1;
          const int a = (code >> 3) & 7;
          const int e = (ix + ox) + (iy + oy) * R->nex + (iz + oz) * R->nex * R->ney;
          const double *ke = ke0 + (size_t)P.elem_mat[R->elem_base + e] * kKe + (3 * a + d) * kNd;
          const int n0 = n + ox + oy * nx + oz * nxy;
#pragma unroll 2
          for (int b = 0; b < 8; b++) {
            int lx, ly, lz;
            local_coords(b, &lx, &ly, &lz);
            const long long mrow = rbase + 3LL * (n0 + lx + ly * nx + lz * nxy);
            const int jx = ix + ox + lx, jy = iy + oy + ly, jz = iz + oz + lz;
            const bool nb_bc = jx == 0 || jx == nx - 1 || jy == 0 || jy == R->ny - 1 || jz == 0 ||
                               jz == R->nz - 1;
            for (int dd = 0; dd < 3; dd++) {
              const double kv = ke[3 * b + dd];
              const double *xv = P.x64 + (mrow + dd) * KS;
#pragma unroll
              for (int k = 0; k < K; k++) r64[k] = fma(kv, xv[k], r64[k]);
              if (need_cold && nb_bc)
#pragma unroll
                for (int k = 0; k < K; k++) rc64[k] = fma(kv, xv[k], rc64[k]);
            }
          }
        }
      }
      TS ri[KS], zi[KS], zero[KS];
      const TS dg = dinv[grow];
#pragma unroll
      for (int k = 0; k < KS; k++) {
        const int kk = k < K ? k : 0;
        const double rk = (k < K) ? -r64[kk] : 0.0;
        ri[k] = (TS)rk;
        zi[k] = dg * ri[k];
        zero[k] = (TS)0;
        if (k < K) {
          v[kk] += rk * rk;
          v[K + kk] += (double)ri[k] * (double)ri[k];
          v[2 * K + kk] += (double)ri[k] * (double)zi[k];
          v[3 * K + kk] += rc64[kk] * rc64[kk];
        }
      }
      store_vec<TS, KS>(rv + grow * KS, ri);
      store_vec<TS, KS>(zv + grow * KS, zi);
      store_vec<TS, KS>(dv + grow * KS, zero);
      store_vec<TS, KS>((TS *)P.p + grow * KS, zero);
      store_vec<TS, KS>((TS *)P.q + grow * KS, zero);
    }
  }
  block_sum<4 * K>(v, lds);
  if (tid == 0)
    for (int i = 0; i < 4 * K; i++) part[i] = v[i];
  if (!arrive_last(&R->cntB, nslab)) return;
  double s4[4 * K];
  group_reduce<4 * K>(P.part + (size_t)R->first_slab * kPart, nslab, kPart, 0, s4, lds);
  if (tid == 0) {
    unsigned done_out = 0;
    for (int k = 0; k < K; k++) {
      const double s[4] = {s4[k], s4[K + k], s4[2 * K + k], s4[3 * K + k]};
      const double rn = sqrt(s[0]);
      R->rnorm[k] = rn;
      if (R->sweep == 0) R->r0[k] = need_cold ? sqrt(s[3]) : rn;
      const bool conv = (k >= R->nload) || R->r0[k] == 0.0 || rn <= sp.rtol * R->r0[k] ||
                        rn < sp.abs_tol || R->sweep >= sp.max_sweeps;
      if (conv) done_out |= 1u << k;
      // absolute target of this sweep: inner_rtol of the sweep's right-hand
      // side, but no tighter than what the outer tolerance still needs
      const double cap = (R->r0[k] > 0.0 && rn > 0.0) ? 0.5 * sp.rtol * R->r0[k] / rn : 0.0;
      R->bnorm_in[k] = fmax(sp.inner_rtol, cap) * sqrt(s[1]);
      R->beta[k] = 0.0;
      R->rz[k] = s[2];
      R->alpha[k] = 0.0;
    }
    const unsigned all = (1u << K) - 1u;
    R->done_out = done_out;
    if ((done_out & all) == all) {
      R->phase = PH_AVG;
    } else {
      R->done_in = done_out;
      R->its = 0;
      R->phase = PH_INNER;
    }
  }
}

// ---------------------------------------------------------------------------
// PH_AVG: volume averages of strain and stress (micro/src/homogenize.c:344-368)
// in FP64, sum_gp (.) w_gp per element through sum_gp B w, then the tangent
// (micro/src/homogenize.c:80).  One load case at a time keeps registers low.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_avg(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[12 * 16];
  __shared__ double bw[144];
  __shared__ double cm[kNmat * 36];
  __shared__ double sig[MM_MAX_RHS][6];
  const int slab = blockIdx.x, tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  if (R->phase != PH_AVG) return;
  const int nslab = R->nslab;
  const long long rbase = R->row_base;
  double *part = P.part + (size_t)slab * kPart;
  const int ls = slab - R->first_slab;
  const int nelm = R->nelm;
  const int epb = (nelm + nslab - 1) / nslab;
  const int e_end = min(nelm, (ls + 1) * epb);
  for (int i = tid; i < 144; i += blockDim.x) bw[i] = P.bw_tab[(size_t)slot * 144 + i];
  for (int i = tid; i < kNmat * 36; i += blockDim.x) cm[i] = P.c_tab[(size_t)slot * kNmat * 36 + i];
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < K; k++) {
    double acc[12];
#pragma unroll
    for (int i = 0; i < 12; i++) acc[i] = 0.0;
    for (int e = ls * epb + tid; e < e_end; e += blockDim.x) {
      const int ex = e % R->nex, ey = (e / R->nex) % R->ney, ez = e / (R->nex * R->ney);
      const int n0 = ex + ey * R->nx + ez * R->nx * R->ny;
      const int mat = P.elem_mat[R->elem_base + e];
      const double *C = cm + mat * 36;
      double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int a = 0; a < 8; a++) {
        int lx, ly, lz;
        local_coords(a, &lx, &ly, &lz);
        const int node = n0 + lx + ly * R->nx + lz * R->nx * R->ny;
        for (int dd = 0; dd < 3; dd++) {
          const double xv = P.x64[(rbase + 3 * node + dd) * KS + k];
#pragma unroll
          for (int vv = 0; vv < 6; vv++) eps[vv] += bw[vv * 24 + 3 * a + dd] * xv;
        }
      }
#pragma unroll
      for (int vv = 0; vv < 6; vv++) {
        double s = 0.0;
        for (int w = 0; w < 6; w++) s += C[vv * 6 + w] * eps[w];
        acc[vv] += eps[vv];
        acc[6 + vv] += s;
      }
    }
    block_sum<12>(acc, lds);
    if (tid == 0)
      for (int i = 0; i < 12; i++) part[k * 12 + i] = acc[i];
  }
  if (!arrive_last(&R->cntA, nslab)) return;
  __shared__ double tot[MM_MAX_RHS][12];
#pragma unroll 1
  for (int k = 0; k < K; k++) {
    double s12[12];
    group_reduce<12>(P.part + (size_t)R->first_slab * kPart, nslab, kPart, k * 12, s12, lds);
    if (tid == 0)
      for (int i = 0; i < 12; i++) tot[k][i] = s12[i];
  }
  if (tid == 0) {
    RveOut *o = P.out + slot;
    double rel = 0.0;
    bool ok = true;
#pragma unroll 1
    for (int k = 0; k < K; k++) {
#pragma unroll 1
      for (int i = 0; i < 12; i++) {
        const double s = tot[k][i] / R->vol;
        if (i < 6) o->strain[k][i] = s;
        else {
          o->stress[k][i - 6] = s;
          sig[k][i - 6] = s;
        }
      }
      if (k < R->nload) {
        const double r0 = R->r0[k], rn = R->rnorm[k];
        rel = fmax(rel, (r0 > 0.0) ? rn / r0 : 0.0);
        if (!(r0 == 0.0 || rn <= sp.rtol * r0 || rn < sp.abs_tol)) ok = false;
      }
    }
    for (int i = 0; i < 36; i++) o->c[i] = 0.0;
    if (R->tangent_mode == MM_TANGENT_UNIT && K >= 6) {
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) o->c[j * 6 + i] = sig[i][j] / R->load[i][i];
    } else if (R->tangent_mode == MM_TANGENT_FD && K >= 7) {
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++)
          o->c[j * 6 + i] = (sig[i + 1][j] - sig[0][j]) / (R->load[i + 1][i] - R->load[0][i]);
    }
    o->residual_rel = rel;
    o->its_total = R->its_total;
    o->sweeps = R->sweep;
    o->status = ok ? MM_OK : MM_NOT_CONVERGED;
    R->status = o->status;
    R->phase = PH_DONE;
    __threadfence_system();
    o->phase = PH_DONE;
    __threadfence_system();
  }
}

// Stand-alone copy of the INNER-phase SpMV (w = A z) over every resident RVE,
// written to q; used only when no RVE is being solved.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_spmv_time(EngPtrs P)
{
  constexpr int KS = vec_stride(K);
  const int slab = blockIdx.x, tid = threadIdx.x, SN = P.SN;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  const RveDev *R = P.rves + slot;
  const int n = P.slab_node0[slab] + tid;
  if (n >= R->nrow / 3) return;
  const long long rbase = R->row_base;
  TS w[3][KS];
  node_spmv<TS, KS>(P.bcols, (const TS *)P.vals, (long long)slab * 27 * SN, (long long)slab * kW * 3 * SN,
                    SN, tid, rbase, (const TS *)P.z, w);
  TS o[3 * KS];
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int k = 0; k < KS; k++) o[d * KS + k] = w[d][k];
  store_vec<TS, 3 * KS>((TS *)P.q + (rbase + 3LL * n) * KS, o);
}

// ---------------------------------------------------------------------------
// launchers (explicit instantiations for the supported load-case counts)

template <typename TS, int K>
static int launch_steps(const EngPtrs &P, const SolveParams &sp, int grid, int nsteps, hipStream_t s)
{
  for (int i = 0; i < nsteps; i++) {
    k_step_a<TS, K><<<grid, P.SN, 0, s>>>(P, sp);
    k_step_b<TS, K><<<grid, P.SN, 0, s>>>(P, sp);
    k_step_res<TS, K><<<grid, P.SN, 0, s>>>(P, sp);
  }
  k_step_avg<TS, K><<<grid, P.SN, 0, s>>>(P, sp);
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename TS, int K>
static int launch_assemble(const EngPtrs &P, int nblk, const int *slab_list, int mode, hipStream_t s)
{
  if (nblk <= 0) return 0;
  k_assemble<TS, K><<<nblk, P.SN, 0, s>>>(nblk, slab_list, P.slab_rve, P.slab_node0, P.rves, P.ke_tab,
                                          P.elem_mat, const_cast<int *>(P.bcols), (TS *)P.vals,
                                          (TS *)P.dinv, P.x64, P.SN, mode);
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename TS, int K>
static int launch_time(const EngPtrs &P, int grid, hipStream_t s)
{
  k_spmv_time<TS, K><<<grid, P.SN, 0, s>>>(P);
  MM_HIP(hipGetLastError());
  return 0;
}

#define MM_DISPATCH(FP64, K, CALL_T)      \
  do {                                    \
    if (FP64) {                           \
      switch (K) {                        \
        case 1: return CALL_T(double, 1); \
        case 6: return CALL_T(double, 6); \
        case 7: return CALL_T(double, 7); \
        default: return MM_ERR_ARG;       \
      }                                   \
    } else {                              \
      switch (K) {                        \
        case 1: return CALL_T(float, 1);  \
        case 6: return CALL_T(float, 6);  \
        case 7: return CALL_T(float, 7);  \
        default: return MM_ERR_ARG;       \
      }                                   \
    }                                     \
  } while (0)

bool eng_supported_rhs(int K) { return K == 1 || K == 6 || K == 7; }

int eng_launch_steps(bool fp64, int K, const EngPtrs &P, const SolveParams &sp, int grid, int nsteps,
                     hipStream_t s)
{
#define CALL_STEPS(T, KK) launch_steps<T, KK>(P, sp, grid, nsteps, s)
  MM_DISPATCH(fp64, K, CALL_STEPS);
#undef CALL_STEPS
}

int eng_launch_assemble(bool fp64, int K, const EngPtrs &P, int nblk, const int *slab_list, int mode,
                        hipStream_t s)
{
#define CALL_ASM(T, KK) launch_assemble<T, KK>(P, nblk, slab_list, mode, s)
  MM_DISPATCH(fp64, K, CALL_ASM);
#undef CALL_ASM
}

int eng_launch_time(bool fp64, int K, const EngPtrs &P, int grid, hipStream_t s)
{
#define CALL_TIME(T, KK) launch_time<T, KK>(P, grid, s)
  MM_DISPATCH(fp64, K, CALL_TIME);
#undef CALL_TIME
}

int eng_launch_admit(int n, const int *slots, const RveDev *states, RveDev *rves, int *slab_rve,
                     int *slab_node0, int SN, hipStream_t s)
{
  if (n <= 0) return 0;
  k_admit<<<n, 256, 0, s>>>(n, slots, states, rves, slab_rve, slab_node0, SN);
  MM_HIP(hipGetLastError());
  return 0;
}

int eng_launch_ke(int n, const int *work, const RveDev *rves, double *ke_tab, double *c_tab_out,
                  double *bw_tab, hipStream_t s)
{
  if (n <= 0) return 0;
  k_ke<<<n, 64, 0, s>>>(n, work, rves, ke_tab, c_tab_out, bw_tab);
  MM_HIP(hipGetLastError());
  return 0;
}

int eng_launch_phase(int nblk, const int2 *work, const RveDev *rves, uint8_t *elem_mat, hipStream_t s)
{
  if (nblk <= 0) return 0;
  k_phase<<<nblk, 256, 0, s>>>(nblk, work, rves, elem_mat);
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mm
