// engine_kernels.hip -- CDNA4 kernels of the RVE engine.
//
// Arena layout in HBM (S = slab rows, W = 81, KS = padded RHS count):
//   vals/cols : [slab][W][S]   column-major ELL slabs, RVE-local column ids
//                              (bit-identical to the reference's cols)
//   vectors   : [row][KS]      x64 (FP64 displacement), d (correction),
//                              r, z, pA/pB (ping-pong search direction), q
//   dinv      : [row]          Jacobi preconditioner
//   part      : [slab][96]     per-block partial sums (FP64)
// Every resident RVE owns a contiguous run of slabs; row_base = first_slab*S.
//
// Kernel A and kernel B advance every resident RVE by one phase step; their
// grids cover all slabs, and each block reads the phase of its RVE.  The
// block that finishes last for an RVE (arrive_last) reduces the partials of
// that RVE in slab order, so results do not depend on which other RVEs share
// the launch.
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

constexpr int kPart = 96;  // doubles of partials per slab

__host__ __device__ constexpr int vec_stride(int K) { return (K == 1) ? 1 : ((K + 1) & ~1); }

// ---------------------------------------------------------------------------
// Admission: scatter new RVE states and their slab ownership.
__global__ void k_admit(int n, const int *__restrict__ slots, const RveDev *__restrict__ states,
                        RveDev *__restrict__ rves, int *__restrict__ slab_rve,
                        int *__restrict__ slab_row0, int S)
{
  const int i = blockIdx.x;
  if (i >= n) return;
  const int slot = slots[i];
  const RveDev &st = states[i];
  if (threadIdx.x == 0) rves[slot] = st;
  for (int s = threadIdx.x; s < st.nslab; s += blockDim.x) {
    slab_rve[st.first_slab + s] = slot;
    slab_row0[st.first_slab + s] = s * S;
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

// ELL assembly: one lane per row.  Writes all W slots of the row (column ids
// and values, -1 / 0 in the tail), the Jacobi diagonal inverse and the
// initial displacement (affine on the boundary, micro/src/homogenize.c:
// 227-242, 335-342).  Each slot value is the sum of the Ke entries of the
// elements sharing the two nodes, in ascending element id, the first one
// stored and the rest added (src/ell.c:46-73).  Boundary rows become identity
// rows; boundary columns are kept (src/ell.c:86-118).
// mode 0: full assembly; mode 1: displacement only (resident matrix reused).
template <typename TV, int K>
__global__ void k_assemble(int nblk, const int *__restrict__ slab_list, const int *__restrict__ slab_rve,
                           const int *__restrict__ slab_row0, const RveDev *__restrict__ rves,
                           const double *__restrict__ ke_tab, const uint8_t *__restrict__ elem_mat,
                           int *__restrict__ cols, TV *__restrict__ vals, TV *__restrict__ dinv,
                           double *__restrict__ x64, int S, int mode)
{
#pragma clang fp contract(off)
  constexpr int KS = vec_stride(K);
  if ((int)blockIdx.x >= nblk) return;
  const int slab = slab_list[blockIdx.x];
  const int slot = slab_rve[slab];
  const RveDev &R = rves[slot];
  const int tid = threadIdx.x;
  const int row = slab_row0[slab] + tid;
  const long long sb = (long long)slab * S * kW;
  const long long grow = R.row_base + row;
  if (row >= R.nrow) {
    if (mode == 0) {
      for (int j = 0; j < kW; j++) {
        cols[sb + (long long)j * S + tid] = -1;
        vals[sb + (long long)j * S + tid] = (TV)0;
      }
      dinv[grow] = (TV)0;
    }
    for (int k = 0; k < KS; k++) x64[grow * KS + k] = 0.0;
    return;
  }
  const int n = row / 3, d = row - 3 * (row / 3);
  const int nx = R.nx, nxy = R.nx * R.ny;
  const int ix = n % nx, iy = (n / nx) % R.ny, iz = n / nxy;
  const int cx = (ix == 0) ? 0 : ((ix == nx - 1) ? 2 : 1);
  const int cy = (iy == 0) ? 0 : ((iy == R.ny - 1) ? 2 : 1);
  const int cz = (iz == 0) ? 0 : ((iz == R.nz - 1) ? 2 : 1);
  const int cls = cx + 3 * cy + 9 * cz;
  const bool bc = (cls != 13);
  if (mode == 0) {
    const int ns = g_slots.nslot[cls];
    const int dslot = g_slots.diag_slot[cls];
    const double *ke0 = ke_tab + (size_t)slot * kNmat * kKe;
    double diag = 0.0;
    for (int t = 0; t < 27; t++) {
      if (t < ns) {
        const int key = g_slots.off[cls][t];
        const int m = n + (key % 3 - 1) + ((key / 3) % 3 - 1) * nx + (key / 9 - 1) * nxy;
        const int nc = g_slots.ncontrib[cls][t];
        double v[3];
        if (bc) {
          for (int dd = 0; dd < 3; dd++) v[dd] = (3 * m + dd == row) ? 1.0 : 0.0;
        } else {
          for (int dd = 0; dd < 3; dd++) v[dd] = 0.0;
          for (int c = 0; c < nc; c++) {
            const int code = g_slots.contrib[cls][t][c];
            const int ox = (code & 1) - 1, oy = ((code >> 1) & 1) - 1, oz = ((code >> 2) & 1) - 1;
            const int a = (code >> 3) & 7, b = (code >> 6) & 7;
            const int e = (ix + ox) + (iy + oy) * R.nex + (iz + oz) * R.nex * R.ney;
            const double *ke = ke0 + (size_t)elem_mat[R.elem_base + e] * kKe + (3 * a + d) * kNd + 3 * b;
            for (int dd = 0; dd < 3; dd++) v[dd] = (c == 0) ? ke[dd] : v[dd] + ke[dd];
          }
        }
        if (t == dslot) diag = v[d];
        for (int dd = 0; dd < 3; dd++) {
          cols[sb + (long long)(3 * t + dd) * S + tid] = 3 * m + dd;
          vals[sb + (long long)(3 * t + dd) * S + tid] = (TV)v[dd];
        }
      } else {
        for (int dd = 0; dd < 3; dd++) {
          cols[sb + (long long)(3 * t + dd) * S + tid] = -1;
          vals[sb + (long long)(3 * t + dd) * S + tid] = (TV)0;
        }
      }
    }
    dinv[grow] = (TV)(1.0 / diag);
  }
  // initial displacement for every load case
  const double X[3] = {ix * R.hx, iy * R.hy, iz * R.hz};
  for (int k = 0; k < KS; k++) {
    double u = 0.0;
    if (k < R.nload && bc) {
      const double *s = R.load[k];
      if (d == 0) u = s[0] * X[0] + s[3] / 2 * X[1] + s[5] / 2 * X[2];
      else if (d == 1) u = s[3] / 2 * X[0] + s[1] * X[1] + s[4] / 2 * X[2];
      else u = s[5] / 2 * X[0] + s[4] / 2 * X[1] + s[2] * X[2];
      x64[grow * KS + k] = u;
    } else if (!(R.warm && k < R.nload)) {
      x64[grow * KS + k] = 0.0;
    }
  }
}

// p = z + beta p_old, evaluated identically wherever it is needed
template <typename TS>
__device__ __forceinline__ TS pform(TS z, TS p, TS b) { return fma(b, p, z); }

template <typename T>
struct Vec2Of;
template <>
struct Vec2Of<float> { using type = float2; };
template <>
struct Vec2Of<double> { using type = double2; };

// Loads KS consecutive values (KS even -> vector loads).
template <typename TS, int KS>
__device__ __forceinline__ void load_vec(const TS *__restrict__ p, TS (&v)[KS])
{
  if constexpr (KS % 2 == 0) {
    using V2 = typename Vec2Of<TS>::type;
    const V2 *q = reinterpret_cast<const V2 *>(p);
#pragma unroll
    for (int k = 0; k < KS / 2; k++) {
      const V2 t = q[k];
      v[2 * k] = t.x;
      v[2 * k + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < KS; k++) v[k] = p[k];
  }
}
template <typename TS, int KS>
__device__ __forceinline__ void store_vec(TS *__restrict__ p, const TS (&v)[KS])
{
  if constexpr (KS % 2 == 0) {
    using V2 = typename Vec2Of<TS>::type;
    V2 *q = reinterpret_cast<V2 *>(p);
#pragma unroll
    for (int k = 0; k < KS / 2; k++) q[k] = V2{v[2 * k], v[2 * k + 1]};
  } else {
#pragma unroll
    for (int k = 0; k < KS; k++) p[k] = v[k];
  }
}

// One ELL row times p = z + beta p_old for KS right-hand sides.
template <typename TS, int KS>
__device__ __forceinline__ void ell_row_spmv(const int *__restrict__ cols, const TS *__restrict__ vals,
                                             long long sb, int S, int tid, long long rbase,
                                             const TS *__restrict__ z, const TS *__restrict__ pold,
                                             const TS (&beta)[KS], TS (&acc)[KS])
{
#pragma unroll
  for (int k = 0; k < KS; k++) acc[k] = (TS)0;
#pragma unroll 3
  for (int j = 0; j < kW; j++) {
    const int c = __builtin_nontemporal_load(&cols[sb + (long long)j * S + tid]);
    const TS v = __builtin_nontemporal_load(&vals[sb + (long long)j * S + tid]);
    if (c >= 0) {
      TS zc[KS], pc[KS];
      load_vec<TS, KS>(z + (rbase + c) * KS, zc);
      load_vec<TS, KS>(pold + (rbase + c) * KS, pc);
#pragma unroll
      for (int k = 0; k < KS; k++) acc[k] = fma(v, pform(zc[k], pc[k], beta[k]), acc[k]);
    }
  }
}


// ---------------------------------------------------------------------------
// Kernel A: PH_INNER (q = A p, alpha) and PH_XUPD (x64 += d).
template <typename TS, int K>
__global__ __launch_bounds__(512) void k_step_a(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[K * 16];
  const int slab = blockIdx.x, tid = threadIdx.x, S = P.S;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  const int phase = R->phase;
  if (phase != PH_INNER && phase != PH_XUPD) return;
  const int row = P.slab_row0[slab] + tid;
  const bool valid = row < R->nrow;
  const long long rbase = R->row_base;
  const long long grow = rbase + row;
  const int nslab = R->nslab;
  double *part = P.part + (size_t)slab * kPart;

  if (phase == PH_INNER) {
    const TS *vals = (const TS *)P.vals;
    const TS *z = (const TS *)P.z;
    const TS *pold = (const TS *)(R->pbuf ? P.pB : P.pA);
    TS *pnew = (TS *)(R->pbuf ? P.pA : P.pB);
    TS *q = (TS *)P.q;
    TS beta[KS];
#pragma unroll
    for (int k = 0; k < KS; k++) beta[k] = (k < K) ? (TS)R->beta[k] : (TS)0;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = 0.0;
    if (valid) {
      TS acc[KS];
      ell_row_spmv<TS, KS>(P.cols, vals, (long long)slab * S * kW, S, tid, rbase, z, pold, beta, acc);
      TS zi[KS], pi[KS];
      load_vec<TS, KS>(z + grow * KS, zi);
      load_vec<TS, KS>(pold + grow * KS, pi);
#pragma unroll
      for (int k = 0; k < KS; k++) pi[k] = pform(zi[k], pi[k], beta[k]);
      store_vec<TS, KS>(pnew + grow * KS, pi);
      store_vec<TS, KS>(q + grow * KS, acc);
#pragma unroll
      for (int k = 0; k < K; k++) v[k] = (double)pi[k] * (double)acc[k];
    }
    block_sum<K>(v, lds);
    if (tid == 0)
      for (int k = 0; k < K; k++) part[k] = v[k];
    if (arrive_last(&R->cntA, nslab) && tid == 0) {
      const double *p0 = P.part + (size_t)R->first_slab * kPart;
      for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int b = 0; b < nslab; b++) s += load_coherent(&p0[(size_t)b * kPart + k]);
        const bool done = (R->done_in >> k) & 1u;
        R->alpha[k] = (done || s == 0.0) ? 0.0 : R->rz[k] / s;
      }
      R->pbuf ^= 1;
    }
    return;
  }

  if (phase == PH_XUPD) {
    if (valid) {
      const TS *d = (const TS *)P.d;
#pragma unroll
      for (int k = 0; k < K; k++) P.x64[grow * KS + k] += (double)d[grow * KS + k];
    }
    if (arrive_last(&R->cntA, nslab) && tid == 0) {
      R->sweep += 1;
      R->phase = PH_RES;
    }
    return;
  }

}

// ---------------------------------------------------------------------------
// PH_AVG: volume averages of strain and stress (micro/src/homogenize.c:344-368)
// in FP64, sum_gp (.) w_gp per element through sum_gp B w, then the tangent
// (micro/src/homogenize.c:80).  One load case at a time keeps registers low.
template <typename TS, int K>
__global__ __launch_bounds__(512) void k_step_avg(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[12 * 16];
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
  __shared__ double bw[144];
  __shared__ double cm[kNmat * 36];
  __shared__ double sig[MM_MAX_RHS][6];
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
  if (arrive_last(&R->cntA, nslab) && tid == 0) {
    const double *p0 = P.part + (size_t)R->first_slab * kPart;
    RveOut *o = P.out + slot;
    double rel = 0.0;
    bool ok = true;
#pragma unroll 1
    for (int k = 0; k < K; k++) {
#pragma unroll 1
      for (int i = 0; i < 12; i++) {
        double s = 0.0;
        for (int b = 0; b < nslab; b++) s += load_coherent(&p0[(size_t)b * kPart + k * 12 + i]);
        s = s / R->vol;
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

// ---------------------------------------------------------------------------
// Kernel B: PH_INNER (d += alpha p, r -= alpha q, z = dinv r, r.r, r.z).
template <typename TS, int K>
__global__ __launch_bounds__(512) void k_step_b(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[2 * K * 16];
  const int slab = blockIdx.x, tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  const int phase = R->phase;
  if (phase != PH_INNER) return;
  const int row = P.slab_row0[slab] + tid;
  const bool valid = row < R->nrow;
  const long long rbase = R->row_base;
  const long long grow = rbase + row;
  const int nslab = R->nslab;
  double *part = P.part + (size_t)slab * kPart;
  TS *dv = (TS *)P.d;
  TS *rv = (TS *)P.r;
  TS *zv = (TS *)P.z;
  const TS *dinv = (const TS *)P.dinv;

  {
    const TS *p = (const TS *)(R->pbuf ? P.pB : P.pA);  // written by kernel A
    const TS *q = (const TS *)P.q;
    double v[2 * K];
#pragma unroll
    for (int i = 0; i < 2 * K; i++) v[i] = 0.0;
    if (valid) {
      TS alpha[KS];
#pragma unroll
      for (int k = 0; k < KS; k++) alpha[k] = (k < K) ? (TS)R->alpha[k] : (TS)0;
      TS di[KS], ri[KS], pi[KS], qi[KS], zi[KS];
      load_vec<TS, KS>(dv + grow * KS, di);
      load_vec<TS, KS>(rv + grow * KS, ri);
      load_vec<TS, KS>(p + grow * KS, pi);
      load_vec<TS, KS>(q + grow * KS, qi);
      const TS dg = dinv[grow];
#pragma unroll
      for (int k = 0; k < KS; k++) {
        di[k] = fma(alpha[k], pi[k], di[k]);
        ri[k] = fma(-alpha[k], qi[k], ri[k]);
        zi[k] = dg * ri[k];
      }
      store_vec<TS, KS>(dv + grow * KS, di);
      store_vec<TS, KS>(rv + grow * KS, ri);
      store_vec<TS, KS>(zv + grow * KS, zi);
#pragma unroll
      for (int k = 0; k < K; k++) {
        v[k] = (double)ri[k] * (double)ri[k];
        v[K + k] = (double)ri[k] * (double)zi[k];
      }
    }
    block_sum<2 * K>(v, lds);
    if (tid == 0)
      for (int i = 0; i < 2 * K; i++) part[K + i] = v[i];
    if (arrive_last(&R->cntB, nslab) && tid == 0) {
      const double *p0 = P.part + (size_t)R->first_slab * kPart;
      R->its += 1;
      unsigned done = R->done_in;
      for (int k = 0; k < K; k++) {
        double rr = 0.0, rz = 0.0;
        for (int b = 0; b < nslab; b++) {
          rr += load_coherent(&p0[(size_t)b * kPart + K + k]);
          rz += load_coherent(&p0[(size_t)b * kPart + 2 * K + k]);
        }
        if ((done >> k) & 1u) {
          R->beta[k] = 0.0;
          continue;
        }
        if (sqrt(rr) <= sp.inner_rtol * R->bnorm_in[k] || R->rz[k] == 0.0) {
          done |= 1u << k;
          R->beta[k] = 0.0;
        } else {
          R->beta[k] = rz / R->rz[k];
          R->rz[k] = rz;
        }
      }
      R->done_in = done;
      const unsigned all = (1u << K) - 1u;
      if ((done & all) == all || R->its >= sp.max_inner_its) {
        R->its_total += R->its;
        R->phase = PH_XUPD;
      }
    }
    return;
  }

}

// ---------------------------------------------------------------------------
// Residual kernel: PH_RES, the FP64 residual r = -K x64 computed matrix-free
// from the element stiffness tables (the Newton / refinement residual of
// micro/src/assembly.c:33-82), then the start of the next PCG sweep.
template <typename TS, int K>
__global__ __launch_bounds__(512) void k_step_res(EngPtrs P, SolveParams sp)
{
  constexpr int KS = vec_stride(K);
  __shared__ double lds[4 * K * 16];
  const int slab = blockIdx.x, tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  if (R->phase != PH_RES) return;
  const int row = P.slab_row0[slab] + tid;
  const bool valid = row < R->nrow;
  const long long rbase = R->row_base;
  const long long grow = rbase + row;
  const int nslab = R->nslab;
  double *part = P.part + (size_t)slab * kPart;
  TS *dv = (TS *)P.d;
  TS *rv = (TS *)P.r;
  TS *zv = (TS *)P.z;
  const TS *dinv = (const TS *)P.dinv;

  // PH_RES.  The relative target refers to the cold-start right-hand side
  // (interior displacement zero, the ||b|| of the first solve); a warm-started
  // RVE measures it at sweep 0 from the boundary neighbours alone.
  const bool need_cold = R->warm && R->sweep == 0;
  double v[4 * K];
#pragma unroll
  for (int i = 0; i < 4 * K; i++) v[i] = 0.0;
  if (valid) {
    const int n = row / 3, d = row - 3 * (row / 3);
    const int nx = R->nx, nxy = R->nx * R->ny;
    const int ix = n % nx, iy = (n / nx) % R->ny, iz = n / nxy;
    const int cx = (ix == 0) ? 0 : ((ix == nx - 1) ? 2 : 1);
    const int cy = (iy == 0) ? 0 : ((iy == R->ny - 1) ? 2 : 1);
    const int cz = (iz == 0) ? 0 : ((iz == R->nz - 1) ? 2 : 1);
    const int cls = cx + 3 * cy + 9 * cz;
    double r64[K], rc64[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      r64[k] = 0.0;
      rc64[k] = 0.0;
    }
    if (cls == 13) {  // boundary rows keep a zero residual (micro/src/assembly.c:59-74)
      const double *ke0 = P.ke_tab + (size_t)slot * kNmat * kKe;
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
      const double rk = (k < K) ? -r64[k < K ? k : 0] : 0.0;
      ri[k] = (TS)rk;
      zi[k] = dg * ri[k];
      zero[k] = (TS)0;
      if (k < K) {
        v[k] = rk * rk;
        v[K + k] = (double)ri[k] * (double)ri[k];
        v[2 * K + k] = (double)ri[k] * (double)zi[k];
        v[3 * K + k] = rc64[k < K ? k : 0] * rc64[k < K ? k : 0];
      }
    }
    store_vec<TS, KS>(rv + grow * KS, ri);
    store_vec<TS, KS>(zv + grow * KS, zi);
    store_vec<TS, KS>(dv + grow * KS, zero);
    store_vec<TS, KS>((TS *)P.pA + grow * KS, zero);
    store_vec<TS, KS>((TS *)P.pB + grow * KS, zero);
  }
  block_sum<4 * K>(v, lds);
  if (tid == 0)
    for (int i = 0; i < 4 * K; i++) part[i] = v[i];
  if (arrive_last(&R->cntB, nslab) && tid == 0) {
    const double *p0 = P.part + (size_t)R->first_slab * kPart;
    unsigned done_out = 0;
    for (int k = 0; k < K; k++) {
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int b = 0; b < nslab; b++)
        for (int i = 0; i < 4; i++) s[i] += load_coherent(&p0[(size_t)b * kPart + i * K + k]);
      const double rn = sqrt(s[0]);
      R->rnorm[k] = rn;
      if (R->sweep == 0) R->r0[k] = need_cold ? sqrt(s[3]) : rn;
      const bool conv = (k >= R->nload) || R->r0[k] == 0.0 || rn <= sp.rtol * R->r0[k] ||
                        rn < sp.abs_tol || R->sweep >= sp.max_sweeps;
      if (conv) done_out |= 1u << k;
      R->bnorm_in[k] = sqrt(s[1]);
      R->rz[k] = s[2];
      R->beta[k] = 0.0;
      R->alpha[k] = 0.0;
    }
    const unsigned all = (1u << K) - 1u;
    R->done_out = done_out;
    if ((done_out & all) == all) {
      R->phase = PH_AVG;
    } else {
      R->done_in = done_out;
      R->its = 0;
      R->pbuf = 0;
      R->phase = PH_INNER;
    }
  }
}

// Stand-alone timing copy of the INNER-phase SpMV for every resident RVE:
// q = A (z + beta p_old), no solver state change.
template <typename TS, int K>
__global__ __launch_bounds__(512) void k_spmv_time(EngPtrs P)
{
  constexpr int KS = vec_stride(K);
  const int slab = blockIdx.x, tid = threadIdx.x, S = P.S;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  const RveDev *R = P.rves + slot;
  const int row = P.slab_row0[slab] + tid;
  if (row >= R->nrow) return;
  const long long rbase = R->row_base;
  TS beta[KS];
#pragma unroll
  for (int k = 0; k < KS; k++) beta[k] = (TS)0.5;
  TS acc[KS];
  ell_row_spmv<TS, KS>(P.cols, (const TS *)P.vals, (long long)slab * S * kW, S, tid, rbase,
                       (const TS *)P.z, (const TS *)P.pA, beta, acc);
  store_vec<TS, KS>((TS *)P.q + (rbase + row) * KS, acc);
}

// ---------------------------------------------------------------------------
// launchers (explicit instantiations for the supported RHS counts)

template <typename TS, int K>
static int launch_steps(const EngPtrs &P, const SolveParams &sp, int grid, int nsteps, hipStream_t s)
{
  for (int i = 0; i < nsteps; i++) {
    k_step_a<TS, K><<<grid, P.S, 0, s>>>(P, sp);
    k_step_b<TS, K><<<grid, P.S, 0, s>>>(P, sp);
    k_step_res<TS, K><<<grid, P.S, 0, s>>>(P, sp);
  }
  k_step_avg<TS, K><<<grid, P.S, 0, s>>>(P, sp);
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename TS, int K>
static int launch_assemble(const EngPtrs &P, int nblk, const int *slab_list, int mode, hipStream_t s)
{
  if (nblk <= 0) return 0;
  k_assemble<TS, K><<<nblk, P.S, 0, s>>>(nblk, slab_list, P.slab_rve, P.slab_row0, P.rves, P.ke_tab,
                                         P.elem_mat, const_cast<int *>(P.cols), (TS *)P.vals,
                                         (TS *)P.dinv, P.x64, P.S, mode);
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename TS, int K>
static int launch_time(const EngPtrs &P, int grid, hipStream_t s)
{
  k_spmv_time<TS, K><<<grid, P.S, 0, s>>>(P);
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
                     int *slab_row0, int S, hipStream_t s)
{
  if (n <= 0) return 0;
  k_admit<<<n, 256, 0, s>>>(n, slots, states, rves, slab_rve, slab_row0, S);
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
