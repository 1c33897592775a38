// engine_kernels.hip -- CDNA4 kernels of the RVE engine.
//
// Arena layout in HBM.  Rows (dofs) are numbered 3*node + d; a slab holds SN
// consecutive nodes (3*SN rows) of one RVE and is processed by one workgroup
// of SN lanes, one lane per node.  A node's 27 neighbour blocks are kept in
// neighbour-key order, key = (dx+1) + 3(dy+1) + 9(dz+1); g_slots.key_slot
// maps a key to the reference's slot of the block in the node's rows.
// Full storage:
//   vals  : [slab][81][3*SN]  value of column 3m + d' of row 3n + d at block
//                             position 3*key + d', column-major inside the
//                             slab so lane i reads its three rows as 12
//                             contiguous bytes
//   bcols : [slab][27][SN]    RVE-local neighbour node of each key (-1 = none)
// Dictionary storage:
//   dict  : [rve][256][27][12] the interior blocks per phase pattern
//   npat  : [slab][SN]        the node's pattern
// Both:
//   vectors [slab][K][3*SN]   x64 (FP64 displacement), d (correction), r, z,
//                             p, q (PCG); K = load cases
//   dinv  : [slab][3*SN]      Jacobi preconditioner
//   part  : [slab][96]        per-workgroup partial sums (FP64)
// Every resident RVE owns a contiguous run of slabs; its first row is
// first_slab * 3 * SN.
//
// The solver kernels advance every resident RVE by one phase step; their grid
// covers all slabs and each workgroup reads the phase of its RVE.  Dot
// products are finished by every workgroup of the next launch from the
// per-slab partials, summed in slab order, so an RVE's results do not depend
// on which other RVEs share the launch, and no workgroup fences for another
// (an agent-scope release in every workgroup costs an L2 write-back each).
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

// LDS scratch of the solver kernels: 16 doubles per reduced value, for the
// largest reduction a kernel performs (4K residual norms, 12 averages).
template <int K>
struct RedLds {
  static constexpr int size = 16 * ((4 * K > 12) ? 4 * K : 12);
};
// The same for the plane march, whose group sums go four values at a time.
template <int K>
struct MarchLds {
  static constexpr int size = 16 * ((K > 4) ? K : 4);
};


// MI355X dispatches workgroup i to XCD i % 8.  The solver grids are a
// multiple of 8 and workgroup i processes slab (i % 8) * (grid / 8) + i / 8,
// so each XCD walks one contiguous run of slabs and the z rows that
// neighbouring slabs of an RVE share stay in that XCD's L2.
constexpr int kXcd = 8;
// Largest grid of kernel R (a multiple of 8; k_step_r loops over the slabs).
constexpr int kRGrid = 2048;
__device__ __forceinline__ int xcd_slab()
{
  const int per = gridDim.x / kXcd;
  return (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
}

// Solver launches run over the list of slabs being solved (act_list[0] of
// them, ascending from act_list[1]); the grid is a power of two >= 64 that
// covers the list.  The list is split in 8 contiguous parts, one per XCD, so
// neighbouring slabs of an RVE share an L2.  Returns false for workgroups
// beyond the list.
__device__ __forceinline__ bool active_in(const int *list, int &slab)
{
  const int n = list[0];
  const int per = (n + kXcd - 1) / kXcd;
  const int j = blockIdx.x / kXcd;
  const int i = (blockIdx.x % kXcd) * per + j;
  if (j >= per || i >= n) return false;
  slab = list[1 + i];
  return true;
}
__device__ __forceinline__ bool active_slab(const EngPtrs &P, int &slab) { return active_in(P.act_list, slab); }

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

// Component d of the affine displacement u = E X of Voigt strain s, shears
// halved (micro/src/homogenize.c:335-342).
__device__ __forceinline__ double affine_u(const double *s, const double (&X)[3], int d)
{
#pragma clang fp contract(off)
  if (d == 0) return s[0] * X[0] + s[3] / 2 * X[1] + s[5] / 2 * X[2];
  if (d == 1) return s[3] / 2 * X[0] + s[1] * X[1] + s[4] / 2 * X[2];
  return s[5] / 2 * X[0] + s[4] / 2 * X[1] + s[2] * X[2];
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
                           uint8_t *__restrict__ npat, double *__restrict__ x64, int SN, int mode)
{
#pragma clang fp contract(off)
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
  const long long xb = (long long)slab * K * RS + 3 * tid;  // node n in the [slab][K][3*SN] vectors
  const int nn = R.nrow / 3;
  if (n >= nn) {
    if (mode == ASM_FULL) {
      for (int t = 0; t < 27; t++) bcols[cb + (long long)t * SN + tid] = -1;
      for (int j = 0; j < kW; j++)
        for (int d = 0; d < 3; d++) vals[vb + (long long)j * RS + 3 * tid + d] = (TV)0;
    }
    if (mode != ASM_X_ONLY) {
      npat[(long long)slab * SN + tid] = 0;
      for (int d = 0; d < 3; d++) dinv[grow0 + d] = (TV)0;
    }
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < K; k++) x64[xb + (long long)k * RS + d] = 0.0;
    return;
  }
  const int nx = R.nx, nxy = R.nx * R.ny;
  const int ix = n % nx, iy = (n / nx) % R.ny, iz = n / nxy;
  const int cls = node_class(ix, iy, iz, nx, R.ny, R.nz);
  const bool bc = (cls != 13);
  if (mode != ASM_X_ONLY) {
    // material pattern of the 8 elements around an interior node: bit c is
    // the phase of the element at offset ((c&1)-1, ((c>>1)&1)-1, (c>>2)-1)
    int pat = 0;
    if (!bc)
      for (int c = 0; c < 8; c++) {
        const int e = (ix + (c & 1) - 1) + (iy + ((c >> 1) & 1) - 1) * R.nex +
                      (iz + (c >> 2) - 1) * R.nex * R.ney;
        pat |= (int)elem_mat[R.elem_base + e] << c;
      }
    npat[(long long)slab * SN + tid] = (uint8_t)pat;
  }
  if (mode == ASM_DICT) {
    // only the diagonal, summed exactly as the full assembly sums it
    double diag[3] = {1.0, 1.0, 1.0};
    if (!bc) {
      const double *ke0 = ke_tab + (size_t)slot * kNmat * kKe;
      const int t = g_slots.diag_slot[13];
      const int nc = g_slots.ncontrib[13][t];
      for (int c = 0; c < nc; c++) {
        const int code = g_slots.contrib[13][t][c];
        const int ox = (code & 1) - 1, oy = ((code >> 1) & 1) - 1, oz = ((code >> 2) & 1) - 1;
        const int a = (code >> 3) & 7, b = (code >> 6) & 7;
        const int e = (ix + ox) + (iy + oy) * R.nex + (iz + oz) * R.nex * R.ney;
        const double *ke = ke0 + (size_t)elem_mat[R.elem_base + e] * kKe + 3 * a * kNd + 3 * b;
        for (int d = 0; d < 3; d++) diag[d] = (c == 0) ? ke[d * kNd + d] : diag[d] + ke[d * kNd + d];
      }
    }
    for (int d = 0; d < 3; d++) dinv[grow0 + d] = (TV)(1.0 / diag[d]);
  }
  if (mode == ASM_FULL) {
    const double *ke0 = ke_tab + (size_t)slot * kNmat * kKe;
    double diag[3] = {0.0, 0.0, 0.0};
    // blocks stored in neighbour-key order; slot t is the reference's position
    for (int key = 0; key < 27; key++) {
      const int t = g_slots.key_slot[cls][key];
      if (t >= 0) {
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
        if (key == 13)
          for (int d = 0; d < 3; d++) diag[d] = v[d][d];
        bcols[cb + (long long)key * SN + tid] = m;
        for (int dd = 0; dd < 3; dd++)
          for (int d = 0; d < 3; d++)
            vals[vb + (long long)(3 * key + dd) * RS + 3 * tid + d] = (TV)v[d][dd];
      } else {
        bcols[cb + (long long)key * SN + tid] = -1;
        for (int dd = 0; dd < 3; dd++)
          for (int d = 0; d < 3; d++) vals[vb + (long long)(3 * key + dd) * RS + 3 * tid + d] = (TV)0;
      }
    }
    for (int d = 0; d < 3; d++) dinv[grow0 + d] = (TV)(1.0 / diag[d]);
  }
  // initial displacement for every load case: affine on the boundary; inside
  // zero, the affine field, or (warm start) the kept displacements combined
  // as sum_j wcoef[k][j] x_old[j] plus the affine field of dload[k]
  const double X[3] = {ix * R.hx, iy * R.hy, iz * R.hz};
  double xo[3][K];
  if (R.warm && !bc)
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < K; k++) xo[d][k] = x64[xb + (long long)k * RS + d];
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < K; k++) {
      if (k >= R.nload) {
        x64[xb + (long long)k * RS + d] = 0.0;
      } else if (bc || (R.affine && !R.warm)) {
        x64[xb + (long long)k * RS + d] = affine_u(R.load[k], X, d);
      } else if (R.warm) {
        double u = affine_u(R.dload[k], X, d);
        for (int j = 0; j < K; j++)
          if (R.wcoef[k][j] != 0.0) u = u + R.wcoef[k][j] * xo[d][j];
        x64[xb + (long long)k * RS + d] = u;
      } else {
        x64[xb + (long long)k * RS + d] = 0.0;
      }
    }
}

// Dictionary of interior ELL blocks per RVE: entry (slot, pattern p) is the
// 3-row block of an interior node whose element at offset code c has phase
// bit c of p.  Each value is the FP64 sum over the contributing elements in
// the order the full assembly (and micro/src/assembly.c:97-109 with
// src/ell.c:46-84) adds them, then rounded to TV, so a dictionary row equals
// the assembled row bit for bit.  Grid: n * 256 workgroups, 324 lanes used.
template <typename TV>
__global__ __launch_bounds__(384) void k_dict(int n, const int *__restrict__ slots,
                                              const double *__restrict__ ke_tab, TV *__restrict__ dict)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x / kDictPat, pat = blockIdx.x % kDictPat;
  if (i >= n) return;
  const int tid = threadIdx.x;
  if (tid >= kDictEntry) return;
  const int slot = slots[i];
  const int key = tid / kDictBlk, dd = (tid % kDictBlk) / 4, d = tid % 4;
  const int t = g_slots.key_slot[13][key];
  double v = 0.0;
  if (d < 3) {
    const double *ke0 = ke_tab + (size_t)slot * kNmat * kKe;
    const int nc = g_slots.ncontrib[13][t];
    for (int c = 0; c < nc; c++) {
      const int code = g_slots.contrib[13][t][c];
      const int a = (code >> 3) & 7, b = (code >> 6) & 7;
      const int mat = (pat >> (code & 7)) & 1;
      const double kv = ke0[(size_t)mat * kKe + (3 * a + d) * kNd + 3 * b + dd];
      v = (c == 0) ? kv : v + kv;
    }
  }
  dict[((size_t)slot * kDictPat + pat) * kDictEntry + tid] = (TV)v;
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

// ---------------------------------------------------------------------------
// Solver vectors are [slab][K][3*SN]: within a slab, each load case is one
// run of the slab's 3*SN rows.  Consecutive lanes (nodes) then read
// consecutive 12-byte triples of one load case, so a wavefront's gather of a
// neighbour block is one contiguous run, and the element-wise passes stream
// each load case's run with a compile-time case index.

// Offset of RVE-local node m's first row in load case 0.
__device__ __forceinline__ long long node_vec(const EngPtrs &P, int first_slab, int m, int K)
{
  return ((long long)(first_slab + (m >> P.sn_shift)) * K) * (3 * P.SN) + 3 * (m & (P.SN - 1));
}

template <typename T, int K>
__device__ __forceinline__ void gather_node(const T *__restrict__ x, long long nb, int RS, T (&v)[3][K])
{
#pragma unroll
  for (int k = 0; k < K; k++)
#pragma unroll
    for (int dd = 0; dd < 3; dd++) v[dd][k] = x[nb + (long long)k * RS + dd];
}

template <typename T, int K>
__device__ __forceinline__ void scatter_node(T *__restrict__ x, long long nb, int RS, const T (&v)[3][K])
{
#pragma unroll
  for (int k = 0; k < K; k++)
#pragma unroll
    for (int d = 0; d < 3; d++) x[nb + (long long)k * RS + d] = v[d][k];
}

// Row block of lane tid (node n) times z, full storage: the 27 neighbour
// blocks in key order, w[d] accumulated over keys and d'.  zc receives the
// node's own z rows.
template <typename TS, int K>
__device__ __forceinline__ void node_spmv(const EngPtrs &P, int first_slab, int slab, int tid, int n,
                                          const TS *__restrict__ z, TS (&w)[3][K], TS (&zc)[3][K])
{
  const int SN = P.SN, RS = 3 * SN;
  const int *__restrict__ bcols = P.bcols;
  const TS *__restrict__ vals = (const TS *)P.vals;
  const long long cb = (long long)slab * 27 * SN, vb = (long long)slab * kW * RS;
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int k = 0; k < K; k++) w[d][k] = (TS)0;
#pragma unroll 3
  for (int key = 0; key < 27; key++) {
    const int m = __builtin_nontemporal_load(&bcols[cb + (long long)key * SN + tid]);
    TS v[3][3];  // [d'][d]
#pragma unroll
    for (int dd = 0; dd < 3; dd++)
#pragma unroll
      for (int d = 0; d < 3; d++)
        v[dd][d] = __builtin_nontemporal_load(&vals[vb + (long long)(3 * key + dd) * RS + 3 * tid + d]);
    if (m >= 0) {
      const long long nb = node_vec(P, first_slab, m, K);
#pragma unroll
      for (int k = 0; k < K; k++) {
        TS x[3];
#pragma unroll
        for (int dd = 0; dd < 3; dd++) x[dd] = z[nb + (long long)k * RS + dd];
#pragma unroll
        for (int dd = 0; dd < 3; dd++)
#pragma unroll
          for (int d = 0; d < 3; d++) w[d][k] = fma(v[dd][d], x[dd], w[d][k]);
      }
    }
  }
  gather_node<TS, K>(z, node_vec(P, first_slab, n, K), RS, zc);
}

// A node's three rows, streamed non-temporally: one 96-bit access in FP32.
// In FP64 a 3-vector access is lowered to 32 bytes (the vector's padded
// size) and would touch the next node's first row, so FP64 uses three 64-bit
// accesses.
typedef float nt_f3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ void nt_load3(const float *p, float (&v)[3])
{
  const nt_f3 t = __builtin_nontemporal_load(reinterpret_cast<const nt_f3 *>(p));
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
}
__device__ __forceinline__ void nt_store3(float *p, const float (&v)[3])
{
  nt_f3 t;
  t.x = v[0];
  t.y = v[1];
  t.z = v[2];
  __builtin_nontemporal_store(t, reinterpret_cast<nt_f3 *>(p));
}
__device__ __forceinline__ void nt_load3(const double *p, double (&v)[3])
{
#pragma unroll
  for (int d = 0; d < 3; d++) v[d] = __builtin_nontemporal_load(p + d);
}
__device__ __forceinline__ void nt_store3(double *p, const double (&v)[3])
{
#pragma unroll
  for (int d = 0; d < 3; d++) __builtin_nontemporal_store(v[d], p + d);
}

// Reads an address that is the same in every lane through the scalar unit.
template <typename T>
__device__ __forceinline__ T uniform_load(const T *p)
{
  return *(const __attribute__((address_space(4))) T *)p;
}

// Raw buffer view of one RVE's run of a solver vector: base at its first
// slab, `bytes` long (below 2^31).  A load takes the lane's byte offset and a
// wave-uniform byte offset (the load case's run) in a scalar register, so the
// per-case address arithmetic of a 64-bit gather disappears.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rve_buffer(const void *base, int bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
typedef float buf_f3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ void buffer_load3(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff,
                                             float (&x)[3])
{
  const buf_f3 v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)off, (int)soff, 0);
  x[0] = v.x;
  x[1] = v.y;
  x[2] = v.z;
}
__device__ __forceinline__ void buffer_load3(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff,
                                             double (&x)[3])
{
#pragma unroll
  for (int d = 0; d < 3; d++)
    x[d] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + 8 * d), (int)soff, 0));
}

// Four values of a 16-byte aligned LDS row in one read.
typedef float lds_f4 __attribute__((ext_vector_type(4)));
typedef double lds_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void lds_row4(const float *p, float (&v)[4])
{
  const lds_f4 t = *(const __attribute__((address_space(3))) lds_f4 *)p;
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
__device__ __forceinline__ void lds_row4(const double *p, double (&v)[4])
{
  const __attribute__((address_space(3))) lds_d2 *q = (const __attribute__((address_space(3))) lds_d2 *)p;
  const lds_d2 a = q[0], b = q[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

// Where a wavefront reads its dictionary blocks from.
enum DictSrc : int {
  DICT_SCALAR = 0,  // one entry for the whole wavefront: scalar unit
  DICT_GLOBAL = 1,  // per lane, through L1
  DICT_LDS = 2      // per lane, from the workgroup's LDS copy
};

// The 27-block product of an interior node from dictionary entry D, the
// neighbours gathered in key order.
template <typename TS, int K, int SRC>
__device__ __forceinline__ void dict_rows(const EngPtrs &P, const TS *D, __amdgpu_buffer_rsrc_t zr, int n, int nx,
                                          int nxy, TS (&w)[3][K], TS (&zc)[3][K])
{
  const int RS = 3 * P.SN;
#pragma unroll 3
  for (int key = 0; key < 27; key++) {
    const int m = n + (key % 3 - 1) + ((key / 3) % 3 - 1) * nx + (key / 9 - 1) * nxy;
    TS v[3][4];  // [d'][d]
    if (SRC == DICT_SCALAR) {
#pragma unroll
      for (int dd = 0; dd < 3; dd++)
#pragma unroll
        for (int d = 0; d < 3; d++) v[dd][d] = uniform_load(D + key * kDictBlk + dd * 4 + d);
    } else if (SRC == DICT_LDS) {
#pragma unroll
      for (int dd = 0; dd < 3; dd++) lds_row4(D + key * kDictBlk + dd * 4, v[dd]);
    } else {
#pragma unroll
      for (int dd = 0; dd < 3; dd++) load_vec<TS, 4>(D + key * kDictBlk + dd * 4, v[dd]);
    }
    // byte offset of node m's load case 0 inside the RVE's run of slabs
    const unsigned off =
        (unsigned)(((m >> P.sn_shift) * K * RS + 3 * (m & (P.SN - 1))) * (int)sizeof(TS));
#pragma unroll
    for (int k = 0; k < K; k++) {
      TS x[3];
      buffer_load3(zr, off, (unsigned)(k * RS * (int)sizeof(TS)), x);
      if (key == 13)
#pragma unroll
        for (int dd = 0; dd < 3; dd++) zc[dd][k] = x[dd];
#pragma unroll
      for (int dd = 0; dd < 3; dd++)
#pragma unroll
        for (int d = 0; d < 3; d++) w[d][k] = fma(v[dd][d], x[dd], w[d][k]);
    }
  }
}

// The dictionary entries a workgroup's nodes use (at most `cap` distinct
// patterns), copied to LDS once per workgroup so that wavefronts whose nodes
// mix patterns read their blocks from LDS instead of through L1.
template <typename TS>
struct alignas(16) DictCache {
  static constexpr int cap = (sizeof(TS) == 4) ? 16 : 8;
  alignas(16) TS v[cap][kDictEntry];
  int slot_of[kDictPat];  // LDS entry of a pattern, -1 if not cached
  int pat_of[cap];
  int count;
};

// Fills the cache; called by every thread of the workgroup.
template <typename TS>
__device__ void dict_cache_fill(const EngPtrs &P, const RveDev *R, int slot, int slab, DictCache<TS> &dc)
{
  const int tid = threadIdx.x;
  for (int i = tid; i < kDictPat; i += blockDim.x) dc.slot_of[i] = 0;
  __syncthreads();
  const int n = P.slab_node0[slab] + tid;
  const int nx = R->nx, nxy = R->nx * R->ny;
  if (n < R->nrow / 3 && node_class(n % nx, (n / nx) % R->ny, n / nxy, nx, R->ny, R->nz) == 13)
    dc.slot_of[P.npat[(long long)slab * P.SN + tid]] = 1;  // same value from every writer
  __syncthreads();
  if (tid < 64) {  // the first wavefront numbers the patterns in use
    int base = 0;
    for (int c = 0; c < kDictPat; c += 64) {
      const bool used = dc.slot_of[c + tid] != 0;
      const unsigned long long m = __ballot(used);
      const int s = base + __popcll(m & ((1ull << tid) - 1ull));
      int v = -1;
      if (used && s < DictCache<TS>::cap) {
        v = s;
        dc.pat_of[s] = c + tid;
      }
      dc.slot_of[c + tid] = v;
      base += __popcll(m);
    }
    if (tid == 0) dc.count = min(base, DictCache<TS>::cap);
  }
  __syncthreads();
  const TS *D0 = (const TS *)P.dict + (size_t)slot * kDictPat * kDictEntry;
  for (int e = tid; e < dc.count * kDictEntry; e += blockDim.x) {
    const int s = e / kDictEntry, off = e - s * kDictEntry;
    dc.v[s][off] = D0[(size_t)dc.pat_of[s] * kDictEntry + off];
  }
  __syncthreads();
}

// The same product from the dictionary: an interior node's 27 blocks come
// from entry npat of its RVE at the structured neighbour offsets; a boundary
// node's row is the identity, whose full-ELL product z + 0 is reproduced
// exactly.  Same fma order as node_spmv (key order), so the same bits.  When every
// interior lane of the wavefront has the same pattern the block values are
// read once per wavefront through the scalar unit; otherwise from the
// workgroup's LDS copy (dc), or through L1 if a pattern did not fit in it.
// zc receives the node's own z rows (the centre block of the gather).
template <typename TS, int K>
__device__ __forceinline__ void node_spmv_dict(const EngPtrs &P, const RveDev *R, int slot, int slab, int tid,
                                               int n, const TS *__restrict__ z, const DictCache<TS> &dc,
                                               TS (&w)[3][K], TS (&zc)[3][K])
{
  const int RS = 3 * P.SN;
  const int nx = R->nx, nxy = R->nx * R->ny;
  const int ix = n % nx, iy = (n / nx) % R->ny, iz = n / nxy;
  const bool bc = node_class(ix, iy, iz, nx, R->ny, R->nz) != 13;
  if (bc) {
    gather_node<TS, K>(z, node_vec(P, R->first_slab, n, K), RS, zc);
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int d = 0; d < 3; d++) w[d][k] = zc[d][k] + (TS)0;
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int k = 0; k < K; k++) {
      w[d][k] = (TS)0;
      zc[d][k] = (TS)0;
    }
  const TS *D0 = (const TS *)P.dict + (size_t)slot * kDictPat * kDictEntry;
  // z of this RVE as a buffer; made here, before the branches below, so that
  // the descriptor is plainly wave-uniform and stays in scalar registers
  const int fs = __builtin_amdgcn_readfirstlane(R->first_slab), ns = __builtin_amdgcn_readfirstlane(R->nslab);
  const __amdgpu_buffer_rsrc_t zr = rve_buffer(z + (long long)fs * K * RS, ns * K * RS * (int)sizeof(TS));
  const int pat = P.npat[(long long)slab * P.SN + tid];
  const int pat0 = __builtin_amdgcn_readfirstlane(pat);
  const int cs = dc.slot_of[pat];
  if (__all(pat == pat0))
    dict_rows<TS, K, DICT_SCALAR>(P, D0 + (size_t)pat0 * kDictEntry, zr, n, nx, nxy, w, zc);
  else if (__all(cs >= 0))
    dict_rows<TS, K, DICT_LDS>(P, &dc.v[cs][0], zr, n, nx, nxy, w, zc);
  else
    dict_rows<TS, K, DICT_GLOBAL>(P, D0 + (size_t)pat * kDictEntry, zr, n, nx, nxy, w, zc);
}

// Either storage; dc is the workgroup's dictionary cache (dictionary storage
// only, filled by dict_cache_fill).
template <typename TS, int K, bool DICT>
__device__ __forceinline__ void slab_node_spmv(const EngPtrs &P, const RveDev *R, int slot, int slab, int tid,
                                               int n, const DictCache<TS> *dc, TS (&w)[3][K], TS (&zc)[3][K])
{
  if (DICT) node_spmv_dict<TS, K>(P, R, slot, slab, tid, n, (const TS *)P.z, *dc, w, zc);
  else node_spmv<TS, K>(P, R->first_slab, slab, tid, n, (const TS *)P.z, w, zc);
}

// The workgroup's dictionary cache for storage DICT (nullptr for full
// storage); every thread of the workgroup calls it.
template <typename TS, bool DICT>
__device__ __forceinline__ const DictCache<TS> *dict_cache(const EngPtrs &P, const RveDev *R, int slot, int slab)
{
  if constexpr (DICT) {
    __shared__ DictCache<TS> dc;
    dict_cache_fill<TS>(P, R, slot, slab, dc);
    return &dc;
  } else {
    return nullptr;
  }
}

// Dot products accumulate with explicit fma, so their rounding is the same in
// every code path that computes them (the two matrix storages give the same
// bits only if it is).
//
// Element-wise passes: lane tid handles the value pairs tid, tid + blockDim,
// ... of each load case's run of `rows` valid rows (a pair is 8 contiguous
// bytes in FP32), the odd last row by lane 0.
template <typename TS>
struct Pair;
template <>
struct Pair<float> {
  using type = float2;
};
template <>
struct Pair<double> {
  using type = double2;
};

// Per-slab partial sums of the launch that produced them, summed over the
// RVE's slabs in slab order (group_sum_cols): every workgroup of the consuming
// launch forms the same totals.  A slab has kPart partial values, value i
// at part[i * part_stride + slab] (value-major, so that a sum over an RVE's
// slabs reads contiguous memory).  Values:
//   [0, K)      p.q            (A -> B)
//   [K, 3K)     r.r, r.z       (B -> A)
//   [3K, 7K)    residual norms (R -> A)
//   [0, 12K)    averages       (R -> A), never pending with the others
struct SlabPart {
  double *p;
  long long s;
  __device__ __forceinline__ double &operator[](int i) const { return p[(size_t)i * s]; }
};
__device__ __forceinline__ SlabPart slab_part(const EngPtrs &P, int slab)
{
  return SlabPart{P.part + slab, P.part_stride};
}

// Start of a refinement sweep from the reduced residual norms s4 =
// {||r64||^2, ||r||^2, r.z, ||r_cold||^2} per load case (FP64 residual r64,
// its rounding r, cold-start right-hand side r_cold).
template <int K>
__device__ void start_sweep(SolveState &N, const double (&s4)[4 * K], const RveDev *R,
                            const SolveParams &sp)
{
  const bool need_cold = (R->warm || R->affine) && N.sweep == 0;
  unsigned done_out = 0;
  for (int k = 0; k < K; k++) {
    const double rn = sqrt(s4[k]);
    N.rnorm[k] = rn;
    if (N.sweep == 0) N.r0[k] = need_cold ? sqrt(s4[3 * K + k]) : rn;
    const bool conv = (k >= R->nload) || N.r0[k] == 0.0 || rn <= sp.rtol * N.r0[k] ||
                      rn < sp.abs_tol || N.sweep >= sp.max_sweeps;
    if (conv) done_out |= 1u << k;
    // absolute target of this sweep: inner_rtol of the sweep's right-hand
    // side, but no tighter than what the outer tolerance still needs
    const double cap = (N.r0[k] > 0.0 && rn > 0.0) ? 0.5 * sp.rtol * N.r0[k] / rn : 0.0;
    N.bnorm_in[k] = fmax(sp.inner_rtol, cap) * sqrt(s4[K + k]);
    N.beta[k] = 0.0;
    N.rz[k] = s4[2 * K + k];
  }
  const unsigned all = (1u << K) - 1u;
  N.done_out = done_out;
  if ((done_out & all) == all) {
    N.phase = PH_AVG;
  } else {
    N.done_in = done_out;
    N.its = 0;
    N.phase = PH_INNER;
  }
}

// End of a PCG iteration from the reduced s2 = {r.r, r.z}: each load case
// stops at its own sweep target; the sweep ends (x += d, then a new FP64
// residual) when all have, or at max_inner_its.  Returns true at sweep end.
template <int K>
__device__ bool inner_check(SolveState &N, const double (&s2)[2 * K], const SolveParams &sp)
{
  N.its += 1;
  unsigned done = N.done_in;
  const unsigned all = (1u << K) - 1u;
  for (int k = 0; k < K; k++) {
    const double rr = s2[k], rz = s2[K + k];
    if ((done >> k) & 1u) {
      N.beta[k] = 0.0;
      continue;
    }
    if (sqrt(rr) <= N.bnorm_in[k] || N.rz[k] == 0.0) {
      done |= 1u << k;
      N.beta[k] = 0.0;
    } else {
      N.beta[k] = rz / N.rz[k];
      N.rz[k] = rz;
    }
  }
  N.done_in = done;
  if ((done & all) == all || N.its >= sp.max_inner_its) {
    N.its_total += N.its;
    N.sweep += 1;
    N.phase = PH_RES;
    return true;
  }
  return false;
}

// Volume averages and tangent of a finished RVE (micro/src/homogenize.c:
// 344-368, :80) from the averaging partials; run by one workgroup.
template <int K, int L>
__device__ void finish_rve(const EngPtrs &P, int slot, const RveDev *R, SolveState &N,
                           const SolveParams &sp, double (&lds)[L])
{
  __shared__ double tot[MM_MAX_RHS][12];
  const double *p0 = slab_part(P, R->first_slab).p;
#pragma unroll 1
  for (int k = 0; k < K; k++) group_sum_cols_lds<12>(p0, R->nslab, P.part_stride, k * 12, &tot[k][0], lds);
  __shared__ double sig[MM_MAX_RHS][6];
  if (threadIdx.x != 0) return;
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
      const double r0 = N.r0[k], rn = N.rnorm[k];
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
  o->its_total = N.its_total;
  o->sweeps = N.sweep;
  o->status = ok ? MM_OK : MM_NOT_CONVERGED;
  N.status = o->status;
  N.phase = PH_DONE;
  __threadfence_system();
  o->phase = PH_DONE;
  __threadfence_system();
}

// The INNER work of kernel A for one slab: w = A z (one lane per node), then
// on the node's own rows p = z + beta p, q = w + beta q (so q = A p without
// gathering p) and the partial p.q.  Every load case is gathered (a per-case
// branch between the gathers would serialize their loads); the cases in
// `skip` have met their sweep target and skip only the update.  The own z
// rows come from the gather, not a second read.
// dc: the workgroup's dictionary cache, filled (dictionary storage only).
template <typename TS, int K, bool DICT>
__device__ __forceinline__ void inner_a_dc(const EngPtrs &P, const RveDev *R, int slot, int slab, unsigned skip,
                                           const double *beta, double (&lds)[RedLds<K>::size],
                                           const DictCache<TS> *dc)
{
  const int tid = threadIdx.x, RS = 3 * P.SN;
  const int n = P.slab_node0[slab] + tid;
  const long long sb = (long long)slab * K * RS;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = 0.0;
  if (n < R->nrow / 3) {
    TS w[3][K], zc[3][K];
    slab_node_spmv<TS, K, DICT>(P, R, slot, slab, tid, n, dc, w, zc);
    TS *__restrict__ p = (TS *)P.p;
    TS *__restrict__ q = (TS *)P.q;
    const long long nb = sb + 3 * tid;
    // every load of the node's p and q rows is issued before the first store
    // (p and q do not alias), so the update costs one memory round trip, not
    // one per row; p and q are streamed non-temporally (next touched a whole
    // working set later), which leaves the cache to z and the dictionaries
    TS pv[3][K], qv[3][K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      if ((skip >> k) & 1u) continue;
      const long long i = nb + (long long)k * RS;
      TS a[3], c[3];
      nt_load3(p + i, a);
      nt_load3(q + i, c);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        pv[d][k] = a[d];
        qv[d][k] = c[d];
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      if ((skip >> k) & 1u) continue;
      const TS b = (TS)beta[k];
      const long long i = nb + (long long)k * RS;
      TS po[3], qo[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        po[d] = fma(b, pv[d][k], zc[d][k]);
        qo[d] = fma(b, qv[d][k], w[d][k]);
        v[k] = fma((double)po[d], (double)qo[d], v[k]);
      }
      nt_store3(p + i, po);
      nt_store3(q + i, qo);
    }
  }
  block_sum<K>(v, lds);
  if (tid == 0) {
    const SlabPart part = slab_part(P, slab);
    for (int k = 0; k < K; k++) part[k] = v[k];
  }
}

template <typename TS, int K, bool DICT>
__device__ __forceinline__ void inner_a(const EngPtrs &P, const RveDev *R, int slot, int slab, unsigned skip,
                                        const double *beta, double (&lds)[RedLds<K>::size])
{
  inner_a_dc<TS, K, DICT>(P, R, slot, slab, skip, beta, lds, dict_cache<TS, DICT>(P, R, slot, slab));
}

// ---------------------------------------------------------------------------
// Kernel A (reads st[2], writes st[0]).
// PH_RESRED: starts the sweep from the residual norms, then continues as
// INNER with beta = 0.  PH_INNER: from the last B's r.r and r.z, either ends
// the sweep (x64 += d, PH_RES) or forms beta and, one lane per node, does
// w = A z, then on the node's own rows p = z + beta p, q = w + beta q (so
// q = A p without gathering p) and the partial p.q.  Load cases that have
// met their sweep target are skipped.  PH_AVGRED: the first slab's workgroup
// writes the result.
template <typename TS, int K, bool DICT>
__global__ __launch_bounds__(256) void k_step_a(EngPtrs P, SolveParams sp)
{
  __shared__ double lds[RedLds<K>::size];
  __shared__ SolveState N;
  __shared__ int s_xupd;
  int slab;
  if (!active_slab(P, slab)) return;
  const int tid = threadIdx.x, SN = P.SN, RS = 3 * SN;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  if (P.a_lds > 0 && R->mcpp > 0) return;  // k_step_a_march's
  const bool writer = slab == R->first_slab;
  const int nslab = R->nslab;
  const int phase0 = R->st[2].phase;
  const double *p0 = slab_part(P, R->first_slab).p;
  if (phase0 == PH_AVGRED) {
    if (!writer) return;
    if (tid == 0) N = R->st[2];
    __syncthreads();
    finish_rve<K>(P, slot, R, N, sp, lds);
    if (tid == 0) R->st[0] = N;
    return;
  }
  if (tid == 0) {
    N = R->st[2];
    s_xupd = 0;
  }
  if (phase0 == PH_RESRED) {
    double s4[4 * K];
    group_sum_cols<4 * K>(p0, nslab, P.part_stride, 3 * K, s4, lds);
    if (tid == 0) start_sweep<K>(N, s4, R, sp);
  } else if (phase0 == PH_INNER) {
    double s2[2 * K];
    if (R->big) {
      for (int i = 0; i < 2 * K; i++) s2[i] = R->rr_rz[i];  // summed by kernel R
    } else {
      group_sum_cols<2 * K>(p0, nslab, P.part_stride, K, s2, lds);
    }
    if (tid == 0) s_xupd = inner_check<K>(N, s2, sp) ? 1 : 0;
  }
  __syncthreads();
  if (writer && tid == 0) R->st[0] = N;
  if (N.phase != PH_INNER && !s_xupd) return;

  const int n0 = P.slab_node0[slab];
  const int rows = min(RS, R->nrow - 3 * n0);  // valid rows of this slab
  const long long sb = (long long)slab * K * RS;

  if (s_xupd) {
    const TS *d = (const TS *)P.d;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const long long b = sb + (long long)k * RS;
      for (int e = tid; e < rows; e += blockDim.x) P.x64[b + e] += (double)d[b + e];
    }
    return;
  }

  // PH_INNER
  inner_a<TS, K, DICT>(P, R, slot, slab, N.done_in, N.beta, lds);
}


// ---------------------------------------------------------------------------
// Kernel A as a plane march.  In the slab kernel every lane gathers its 27
// neighbours' z through the vector memory path, 27 accesses per node and load
// case, which bounds the kernel by address processing rather than by HBM.
// Here the workgroup of tile t takes an in-plane chunk of SN nodes (in-plane
// index j = c SN + lane, node z nxy + j) through a run of mL z-planes and
// keeps, for three consecutive planes, the window of z around the chunk
// (in-plane indices c SN - nx - 1 .. c SN + SN + nx: every in-plane neighbour
// of the chunk) in an LDS ring, one row-major run per plane, load case and
// direction.  A node's 27 neighbours are then LDS reads at fixed offsets, and
// z is read from memory about once per launch.  The plane two ahead is loaded
// into registers while the current plane is computed and stored into the
// ring slot that the plane behind frees.  Load cases go in groups of
// march_cases(sizeof(TS)) to bound the ring, each group a tile of its own
// (more, shorter tiles: a lightly loaded GPU still fills), and a tile whose
// cases have all met their sweep target returns at once.  w, the p/q update
// and the partial p.q of a node are inner_a's arithmetic (the SpMV in the
// same fma order, so the same bits); block b's p.q partials of the group's
// cases go to slab first_slab + b and the RVE's slabs beyond its blocks
// hold 0.  RVEs with mcpp == 0 (a window wider than
// two positions per lane) stay with k_step_a, which skips the others.

struct MarchGeo {
  int nx, ny, nz, nxy, W;   // RVE nodes per axis and per plane; window length
  int z0, z1;               // the tile's planes
  int wb;                   // in-plane index of window position 0
  int j, pos0;              // the lane's in-plane node and its window position
  bool lane_ok, plane_int;  // j is in the chunk; j is off the four side faces
};

// Tile t = chunk c + mcpp * run r: the chunks split a plane's nxy nodes and
// the runs the nz planes as evenly as integers allow (at most SN nodes and
// mL planes each), so that the tiles of an RVE cost the same.
__device__ __forceinline__ MarchGeo march_geo(const EngPtrs &P, const RveDev *R, int t)
{
  MarchGeo g;
  g.nx = R->nx;
  g.ny = R->ny;
  g.nz = R->nz;
  g.nxy = g.nx * g.ny;
  const int c = t % R->mcpp, zr = t / R->mcpp;
  const int c0 = c * g.nxy / R->mcpp, c1 = (c + 1) * g.nxy / R->mcpp;
  g.z0 = zr * g.nz / R->mnzc;
  g.z1 = (zr + 1) * g.nz / R->mnzc;
  g.W = (c1 - c0) + 2 * g.nx + 2;
  g.wb = c0 - g.nx - 1;
  g.j = c0 + (int)threadIdx.x;
  g.lane_ok = g.j < c1;
  g.pos0 = (g.lane_ok ? (int)threadIdx.x : 0) + g.nx + 1;  // idle lanes read a valid position
  const int ix = g.j % g.nx, iy = g.j / g.nx;
  g.plane_int = g.lane_ok && ix > 0 && ix < g.nx - 1 && iy > 0 && iy < g.ny - 1;
  return g;
}

// Registers <- plane zz of load cases g0 .. g0+GR-1 at the lane's window
// positions lane and lane + SN (the window is at most 2 SN long).  Positions
// off the RVE's plane, and planes outside the RVE, are 0: no node reads them.
template <typename TS, int K, int GR>
__device__ __forceinline__ void march_fetch(const EngPtrs &P, __amdgpu_buffer_rsrc_t zr, const MarchGeo &g, int zz,
                                            int g0, TS (&pf)[2][GR][3])
{
  const int RS = 3 * P.SN;
#pragma unroll
  for (int i = 0; i < 2; i++) {
#pragma unroll
    for (int gk = 0; gk < GR; gk++)
#pragma unroll
      for (int d = 0; d < 3; d++) pf[i][gk][d] = (TS)0;
    const int pos = (int)threadIdx.x + i * P.SN, jj = g.wb + pos;
    if (zz >= 0 && zz < g.nz && pos < g.W && jj >= 0 && jj < g.nxy) {
      const int m = zz * g.nxy + jj;
      const unsigned off = (unsigned)(((m >> P.sn_shift) * K * RS + 3 * (m & (P.SN - 1))) * (int)sizeof(TS));
#pragma unroll
      for (int gk = 0; gk < GR; gk++)
        buffer_load3(zr, off, (unsigned)((g0 + gk) * RS * (int)sizeof(TS)), pf[i][gk]);
    }
  }
}

// Ring slot s <- the registers of march_fetch.
template <typename TS, int GR>
__device__ __forceinline__ void march_put(TS *ring, const MarchGeo &g, int s, const TS (&pf)[2][GR][3])
{
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const int pos = (int)threadIdx.x + i * (int)blockDim.x;
    if (pos < g.W)
#pragma unroll
      for (int gk = 0; gk < GR; gk++)
#pragma unroll
        for (int d = 0; d < 3; d++) ring[((s * GR + gk) * 3 + d) * g.W + pos] = pf[i][gk][d];
  }
}

// The dictionary entries of the patterns a tile's interior nodes use, the
// first kMarchDictCap of them in pattern order, copied to LDS (dv) once per
// tile; slot_of[p] is the LDS entry of pattern p or -1.  Run by every thread.
template <typename TS>
__device__ void march_dict_fill(const EngPtrs &P, const RveDev *R, int slot, int fs, const MarchGeo &g, TS *dv,
                                signed char *slot_of)
{
  __shared__ int s_pat[kMarchDictCap];
  __shared__ int s_cnt;
  const int tid = threadIdx.x;
  for (int i = tid; i < kDictPat; i += blockDim.x) slot_of[i] = 0;
  __syncthreads();
  // the pattern bytes of every plane are loaded before the first is used
  // (one memory round trip, not one per plane)
  if (g.plane_int) {
    const int za = max(g.z0, 1), zb = min(g.z1, g.nz - 1);
#pragma unroll 8
    for (int z = za; z < zb; z++) {
      const int n = z * g.nxy + g.j;
      slot_of[P.npat[(long long)(fs + (n >> P.sn_shift)) * P.SN + (n & (P.SN - 1))]] = 1;  // same value from every writer
    }
  }
  __syncthreads();
  if (tid < 64) {  // the first wavefront numbers the patterns in use
    int base = 0;
    for (int c = 0; c < kDictPat; c += 64) {
      const bool used = slot_of[c + tid] != 0;
      const unsigned long long m = __ballot(used);
      const int sidx = base + __popcll(m & ((1ull << tid) - 1ull));
      int v = -1;
      if (used && sidx < kMarchDictCap) {
        v = sidx;
        s_pat[sidx] = c + tid;
      }
      slot_of[c + tid] = (signed char)v;
      base += __popcll(m);
    }
    if (tid == 0) s_cnt = min(base, kMarchDictCap);
  }
  __syncthreads();
  // entries in 16-byte pieces (an entry is 81 or 162 of them), four loads in
  // flight per lane before the first store
  constexpr int kPieces = kDictEntry * (int)sizeof(TS) / 16;
  const uint4 *D0 = reinterpret_cast<const uint4 *>((const TS *)P.dict + (size_t)slot * kDictPat * kDictEntry);
  uint4 *dq = reinterpret_cast<uint4 *>(dv);
  const int total = s_cnt * kPieces;
  for (int e0 = tid; e0 < total; e0 += 4 * (int)blockDim.x) {
    // pieces past the end load the last one again (no branch between the
    // loads, so they stay in registers and in flight together)
    uint4 t[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = min(e0 + u * (int)blockDim.x, total - 1);
      const int sidx = e / kPieces;
      t[u] = D0[(size_t)s_pat[sidx] * kPieces + (e - sidx * kPieces)];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e < total) dq[e] = t[u];
    }
  }
  __syncthreads();
}

// w += the 27-block product of an interior node from dictionary entry D (a
// wave-uniform global address, the tile's LDS copy, or a per-lane global
// address: SRC), the neighbours read from the ring in key order.
template <typename TS, int GR, int SRC>
__device__ __forceinline__ void march_rows(const TS *D, const TS *ring, const MarchGeo &g, int z, TS (&w)[3][GR])
{
  const int W = g.W;
#pragma unroll 1
  for (int dz = 0; dz < 3; dz++) {
    const int s = (z + dz + 2) % 3;  // ring slot of plane z + dz - 1
#pragma unroll 1
    for (int dy = 0; dy < 3; dy++) {
      const TS *rb = ring + s * GR * 3 * W + g.pos0 + (dy - 1) * g.nx - 1;
      const TS *Db = D + (3 * dy + 9 * dz) * kDictBlk;
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        TS v[3][4];  // [d'][d]
        if (SRC == DICT_SCALAR) {
#pragma unroll
          for (int dd = 0; dd < 3; dd++)
#pragma unroll
            for (int d = 0; d < 3; d++) v[dd][d] = uniform_load(Db + dx * kDictBlk + dd * 4 + d);
        } else if (SRC == DICT_LDS) {
#pragma unroll
          for (int dd = 0; dd < 3; dd++) lds_row4(Db + dx * kDictBlk + dd * 4, v[dd]);
        } else {
#pragma unroll
          for (int dd = 0; dd < 3; dd++) load_vec<TS, 4>(Db + dx * kDictBlk + dd * 4, v[dd]);
        }
#pragma unroll
        for (int gk = 0; gk < GR; gk++) {
          TS x[3];
#pragma unroll
          for (int dd = 0; dd < 3; dd++) x[dd] = rb[(gk * 3 + dd) * W + dx];
#pragma unroll
          for (int dd = 0; dd < 3; dd++)
#pragma unroll
            for (int d = 0; d < 3; d++) w[d][gk] = fma(v[dd][d], x[dd], w[d][gk]);
        }
      }
    }
  }
}

// Pattern byte of the lane's node in plane z (0 off the interior).
__device__ __forceinline__ int march_pat(const EngPtrs &P, const MarchGeo &g, int fs, int z)
{
  if (!(g.plane_int && z > 0 && z < g.nz - 1)) return 0;
  const int n = z * g.nxy + g.j;
  return P.npat[(long long)(fs + (n >> P.sn_shift)) * P.SN + (n & (P.SN - 1))];
}

// One plane z of the tile for load cases g0 .. g0+GR-1: w = A z from the
// ring, then the node's p/q update and its p.q into v (as inner_a).  pat is
// the node's pattern byte (dictionary storage), loaded one plane ahead: the
// load of the next plane's byte is issued after this plane's p/q loads, so
// that waiting for it never waits for them (vector loads retire in order).
template <typename TS, int K, int GR, bool DICT>
__device__ __forceinline__ void march_plane(const EngPtrs &P, int slot, int fs, const MarchGeo &g, const TS *ring,
                                            const TS *dv, const signed char *slot_of, int z, int pat, int &pat_next,
                                            int g0, unsigned skip, const double *beta, double (&v)[K])
{
  const int SN = P.SN, RS = 3 * SN, W = g.W;
  const int n = z * g.nxy + g.j;
  const int sl = fs + (n >> P.sn_shift), tl = n & (SN - 1);  // the node's slab and lane there
  const bool interior = g.plane_int && z > 0 && z < g.nz - 1;
  const long long nb = (long long)sl * K * RS + 3 * tl;
  TS *__restrict__ p = (TS *)P.p;
  TS *__restrict__ q = (TS *)P.q;
  // the node's p and q rows, loaded first so that they arrive during the SpMV
  TS pv[3][GR], qv[3][GR];
  if (g.lane_ok)
#pragma unroll
    for (int gk = 0; gk < GR; gk++) {
      if ((skip >> (g0 + gk)) & 1u) continue;
      const long long i = nb + (long long)(g0 + gk) * RS;
      TS a[3], c[3];
      nt_load3(p + i, a);
      nt_load3(q + i, c);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        pv[d][gk] = a[d];
        qv[d][gk] = c[d];
      }
    }
  if (DICT && z + 1 < g.z1) pat_next = march_pat(P, g, fs, z + 1);
  TS w[3][GR], zc[3][GR];
  const int sc = z % 3;
#pragma unroll
  for (int gk = 0; gk < GR; gk++)
#pragma unroll
    for (int d = 0; d < 3; d++) {
      zc[d][gk] = ring[((sc * GR + gk) * 3 + d) * W + g.pos0];
      w[d][gk] = (TS)0;
    }
  if constexpr (DICT) {
    // interior lanes: their dictionary entry, through the scalar unit when
    // the wavefront's interior lanes share one pattern, else from the tile's
    // LDS copy, else (a pattern beyond the copy) per lane through L1
    const unsigned long long im = __ballot(interior);
    if (im != 0) {
      const TS *D0 = (const TS *)P.dict + (size_t)slot * kDictPat * kDictEntry;
      const int pat0 = __builtin_amdgcn_readlane(pat, (int)__builtin_ctzll(im));
      const int cs = interior ? (int)slot_of[pat] : 0;
      if (__all(!interior || pat == pat0)) {
        if (interior) march_rows<TS, GR, DICT_SCALAR>(D0 + (size_t)pat0 * kDictEntry, ring, g, z, w);
      } else if (__all(!interior || cs >= 0)) {
        if (interior) march_rows<TS, GR, DICT_LDS>(dv + cs * kDictEntry, ring, g, z, w);
      } else {
        if (interior) march_rows<TS, GR, DICT_GLOBAL>(D0 + (size_t)pat * kDictEntry, ring, g, z, w);
      }
    }
    if (!interior)
#pragma unroll
      for (int gk = 0; gk < GR; gk++)
#pragma unroll
        for (int d = 0; d < 3; d++) w[d][gk] = zc[d][gk] + (TS)0;
  } else if (g.lane_ok) {
    // full storage: the node's block columns and values; a stored neighbour
    // is the structured one of its key, which sits in the ring
    const int *__restrict__ bc = P.bcols + (long long)sl * 27 * SN + tl;
    const TS *__restrict__ vr = (const TS *)P.vals + (long long)sl * kW * RS + 3 * tl;
#pragma unroll 1
    for (int dz = 0; dz < 3; dz++) {
      const int s = (z + dz + 2) % 3;
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
        const int pb = g.pos0 + (dy - 1) * g.nx - 1;
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
          const int key = dx + 3 * dy + 9 * dz;
          const int m = __builtin_nontemporal_load(bc + (long long)key * SN);
          TS vb[3][3];
#pragma unroll
          for (int dd = 0; dd < 3; dd++)
#pragma unroll
            for (int d = 0; d < 3; d++) vb[dd][d] = __builtin_nontemporal_load(vr + (long long)(3 * key + dd) * RS + d);
          if (m >= 0)
#pragma unroll
            for (int gk = 0; gk < GR; gk++) {
              TS x[3];
#pragma unroll
              for (int dd = 0; dd < 3; dd++) x[dd] = ring[((s * GR + gk) * 3 + dd) * W + pb + dx];
#pragma unroll
              for (int dd = 0; dd < 3; dd++)
#pragma unroll
                for (int d = 0; d < 3; d++) w[d][gk] = fma(vb[dd][d], x[dd], w[d][gk]);
            }
        }
      }
    }
  }
  if (!g.lane_ok) return;
#pragma unroll
  for (int gk = 0; gk < GR; gk++) {
    const int k = g0 + gk;
    if ((skip >> k) & 1u) continue;
    const TS b = (TS)beta[k];
    const long long i = nb + (long long)k * RS;
    TS po[3], qo[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      po[d] = fma(b, pv[d][gk], zc[d][gk]);
      qo[d] = fma(b, qv[d][gk], w[d][gk]);
      v[k] = fma((double)po[d], (double)qo[d], v[k]);
    }
    nt_store3(p + i, po);
    nt_store3(q + i, qo);
  }
}

// One round of load cases g0 .. g0+GR-1 over the tile's planes.  Every lane
// of the workgroup runs it (barriers inside).  The prologue's three planes
// and the first pattern bytes are loaded before the dictionary cache is
// filled (dictionary storage), so that their memory latency overlaps it.
template <typename TS, int K, int GR, bool DICT>
__device__ __forceinline__ void march_round(const EngPtrs &P, const RveDev *R, int slot, int fs, const MarchGeo &g,
                                            TS *ring, TS *dv, signed char *slot_of, __amdgpu_buffer_rsrc_t zr,
                                            int g0, unsigned skip, const double *beta, double (&v)[K])
{
  // prologue: planes z0 - 1, z0, z0 + 1 (slot = plane % 3; plane -1 unused)
  TS p0[2][GR][3], p1[2][GR][3], pf[2][GR][3];
  march_fetch<TS, K, GR>(P, zr, g, g.z0 - 1, g0, p0);
  march_fetch<TS, K, GR>(P, zr, g, g.z0, g0, p1);
  march_fetch<TS, K, GR>(P, zr, g, g.z0 + 1, g0, pf);
  int pat = DICT ? march_pat(P, g, fs, g.z0) : 0;
  if constexpr (DICT) march_dict_fill<TS>(P, R, slot, fs, g, dv, slot_of);
  march_put<TS, GR>(ring, g, (g.z0 + 2) % 3, p0);
  march_put<TS, GR>(ring, g, g.z0 % 3, p1);
  march_put<TS, GR>(ring, g, (g.z0 + 1) % 3, pf);
  __syncthreads();
  const int zlast = min(g.z1, g.nz - 1);  // the last plane the tile reads
#pragma unroll 1
  for (int z = g.z0; z < g.z1; z++) {
    const bool pre = z + 2 <= zlast;
    if (pre) march_fetch<TS, K, GR>(P, zr, g, z + 2, g0, pf);
    int pat_next = 0;
    march_plane<TS, K, GR, DICT>(P, slot, fs, g, ring, dv, slot_of, z, pat, pat_next, g0, skip, beta, v);
    pat = pat_next;
    if (pre) {
      __syncthreads();  // plane z - 1 read by every lane
      march_put<TS, GR>(ring, g, (z + 2) % 3, pf);
      __syncthreads();
    }
  }
  __syncthreads();  // the ring is free for the next round
}

// The round of case group cg (cases G0 .. G0+GR-1 for the group whose G0 is
// cg * march_cases).
template <typename TS, int K, bool DICT, int G0>
__device__ __forceinline__ void march_group(const EngPtrs &P, const RveDev *R, int slot, int fs, const MarchGeo &g,
                                            TS *ring, TS *dv, signed char *slot_of, __amdgpu_buffer_rsrc_t zr,
                                            int cg, unsigned skip, const double *beta, double (&v)[K])
{
  constexpr int GM = march_cases((int)sizeof(TS));
  constexpr int GR = (K - G0 < GM) ? K - G0 : GM;
  if (cg == G0 / GM) march_round<TS, K, GR, DICT>(P, R, slot, fs, g, ring, dv, slot_of, zr, G0, skip, beta, v);
  if constexpr (G0 + GR < K)
    march_group<TS, K, DICT, G0 + GR>(P, R, slot, fs, g, ring, dv, slot_of, zr, cg, skip, beta, v);
}

extern __shared__ __attribute__((aligned(16))) double march_dyn[];

// The INNER work of tile t (block t % mnt, case group t / mnt) of an RVE
// with mcpp > 0; the dictionary cache and the LDS ring are the launch's
// dynamic LDS.
template <typename TS, int K, bool DICT>
__device__ void march_tile(const EngPtrs &P, const RveDev *R, int slot, int t, unsigned skip, const double *beta,
                           double (&lds)[MarchLds<K>::size])
{
  constexpr int GM = march_cases((int)sizeof(TS));
  const int b = t % R->mnt, cg = t / R->mnt;
  const int k0 = cg * GM, k1 = min(K, k0 + GM);
  const unsigned gmask = ((1u << k1) - 1u) & ~((1u << k0) - 1u);
  if (t == 0) {  // the RVE's slabs without a block contribute 0 to p.q
    const int rest = R->nslab - R->mnt;
    for (int i = threadIdx.x; i < rest * K; i += blockDim.x) {
      const int k = i / rest, sb = R->mnt + i % rest;
      slab_part(P, R->first_slab + sb)[k] = 0.0;
    }
  }
  const int slab = R->first_slab + b;
  if ((skip & gmask) == gmask) {  // every case of the group is done
    if (threadIdx.x == 0)
      for (int k = k0; k < k1; k++) slab_part(P, slab)[k] = 0.0;
    return;
  }
  const int RS = 3 * P.SN;
  const MarchGeo g = march_geo(P, R, b);
  const int fs = __builtin_amdgcn_readfirstlane(R->first_slab), ns = __builtin_amdgcn_readfirstlane(R->nslab);
  const __amdgpu_buffer_rsrc_t zr =
      rve_buffer((const TS *)P.z + (long long)fs * K * RS, ns * K * RS * (int)sizeof(TS));
  // dynamic LDS: the dictionary cache (dictionary storage), then the ring
  unsigned char *lds0 = reinterpret_cast<unsigned char *>(march_dyn);
  TS *dv = reinterpret_cast<TS *>(lds0);
  signed char *slot_of = reinterpret_cast<signed char *>(lds0 + kMarchDictCap * kDictEntry * sizeof(TS));
  TS *ring = reinterpret_cast<TS *>(lds0 + (DICT ? march_dict_bytes((int)sizeof(TS)) : 0));
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = 0.0;
  march_group<TS, K, DICT, 0>(P, R, slot, fs, g, ring, dv, slot_of, zr, cg, skip, beta, v);
  block_sum<K>(v, lds);
  if (threadIdx.x == 0) {
    const SlabPart part = slab_part(P, slab);
    for (int k = k0; k < k1; k++) part[k] = v[k];
  }
}

// Kernel A over the tile list act_tile (the phase handling of k_step_a; the
// workgroup of tile 0 writes the RVE's state, and the x64 += d of a sweep end
// covers slabs t, t + mnt, ...).
template <typename TS, int K, bool DICT>
__device__ __forceinline__ void step_a_march(const EngPtrs &P, const SolveParams &sp)
{
  __shared__ double lds[MarchLds<K>::size];
  __shared__ SolveState N;
  __shared__ int s_xupd;
  int slab;
  if (!active_in(P.act_tile, slab)) return;
  const int tid = threadIdx.x, RS = 3 * P.SN;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  const int t = slab - R->first_slab;
  if (R->mcpp == 0) return;  // not listed: k_step_a's
  const bool writer = t == 0;
  const int nslab = R->nslab;
  const int phase0 = R->st[2].phase;
  const double *p0 = slab_part(P, R->first_slab).p;
  if (phase0 == PH_AVGRED) {
    if (!writer) return;
    if (tid == 0) N = R->st[2];
    __syncthreads();
    finish_rve<K>(P, slot, R, N, sp, lds);
    if (tid == 0) R->st[0] = N;
    return;
  }
  if (tid == 0) {
    N = R->st[2];
    s_xupd = 0;
  }
  // the totals of group_sum_cols; the sweep start's 4K values four at a time
  // (fewer live registers), the every-iteration 2K in one pass (one memory
  // round trip per tile)
  if (phase0 == PH_RESRED) {
    __shared__ double tot[4 * K];
    group_sum_cols_lds<4 * K>(p0, nslab, P.part_stride, 3 * K, tot, lds);
    if (tid == 0) start_sweep<K>(N, *reinterpret_cast<const double(*)[4 * K]>(tot), R, sp);
  } else if (phase0 == PH_INNER) {
    // scratch: the dynamic LDS, which the march has not started to use
    double(&plds)[16 * 2 * K] = *reinterpret_cast<double(*)[16 * 2 * K]>(march_dyn);
    double s2[2 * K];
    if (R->big) {
      for (int i = 0; i < 2 * K; i++) s2[i] = R->rr_rz[i];  // summed by kernel R
    } else {
      group_sum_cols<2 * K>(p0, nslab, P.part_stride, K, s2, plds);
    }
    if (tid == 0) s_xupd = inner_check<K>(N, s2, sp) ? 1 : 0;
  }
  __syncthreads();
  if (writer && tid == 0) R->st[0] = N;
  if (N.phase != PH_INNER && !s_xupd) return;

  if (s_xupd) {
    const TS *d = (const TS *)P.d;
    for (int b = t; b < nslab; b += R->mnt * R->mncg) {
      const int sl = R->first_slab + b;
      const int rows = min(RS, R->nrow - 3 * P.slab_node0[sl]);
      const long long sb = (long long)sl * K * RS;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const long long o = sb + (long long)k * RS;
        for (int e = tid; e < rows; e += blockDim.x) P.x64[o + e] += (double)d[o + e];
      }
    }
    return;
  }

  // PH_INNER
  march_tile<TS, K, DICT>(P, R, slot, t, N.done_in, N.beta, lds);
}

// Dictionary storage: bounded to 128 VGPRs (four waves per SIMD, what the
// LDS of four workgroups per CU allows) without spills.  Full storage keeps
// its block values in registers and is left unbounded.
template <typename TS, int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_step_a_march_dict(EngPtrs P,
                                                                                                  SolveParams sp)
{
  step_a_march<TS, K, true>(P, sp);
}
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_a_march_full(EngPtrs P, SolveParams sp)
{
  step_a_march<TS, K, false>(P, sp);
}

// ---------------------------------------------------------------------------
// Kernel B (reads st[0], writes st[1]).  PH_INNER: alpha = r.z / p.q from
// the last A's partials, then over the slab's values d += alpha p,
// z -= alpha dinv q and the partial r.r, r.z, for the load cases that are
// still iterating.  The residual itself is not stored: z = dinv r carries it
// and r = z / dinv is formed in registers for the dot products.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_b(EngPtrs P, SolveParams sp)
{
  using T2 = typename Pair<TS>::type;
  __shared__ double lds[2 * K * 16];
  __shared__ double s_alpha[K];
  int slab;
  if (!active_slab(P, slab)) return;
  const int tid = threadIdx.x, RS = 3 * P.SN;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  RveDev *R = P.rves + slot;
  const SolveState *S = &R->st[0];
  if (slab == R->first_slab && tid == 0) R->st[1] = *S;
  if (S->phase != PH_INNER) return;
  const unsigned skip = S->done_in;
  const int nslab = R->nslab;
  double s1[K];
  if (R->big) {
    for (int k = 0; k < K; k++) s1[k] = R->pq[k];  // summed by kernel S
  } else {
    group_sum_cols<K>(slab_part(P, R->first_slab).p, nslab, P.part_stride, 0, s1, lds);
  }
  if (tid == 0)
    for (int k = 0; k < K; k++)
      s_alpha[k] = (((skip >> k) & 1u) || s1[k] == 0.0) ? 0.0 : S->rz[k] / s1[k];
  __syncthreads();
  const int n0 = P.slab_node0[slab];
  const int rows = min(RS, R->nrow - 3 * n0);
  const int npair = rows >> 1;
  const long long sb = (long long)slab * K * RS;
  const TS *dinv = (const TS *)P.dinv + (long long)slab * RS;
  TS *__restrict__ dv = (TS *)P.d;
  TS *__restrict__ zv = (TS *)P.z;
  const TS *__restrict__ p = (const TS *)P.p;
  const TS *__restrict__ q = (const TS *)P.q;
  double vu[2 * K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    vu[k] = 0.0;
    vu[K + k] = 0.0;
    if ((skip >> k) & 1u) continue;
    const TS a = (TS)s_alpha[k];
    const long long kb = sb + (long long)k * RS;
    // two pairs per trip, both pairs' loads issued before either's stores;
    // the dots still take the pairs in order tid, tid + blockDim, ...
    const int bs = blockDim.x;
    for (int g = tid; g < npair; g += 2 * bs) {
      const bool two = g + bs < npair;
      T2 dg[2], di[2], zi[2], pi[2], qi[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (h == 1 && !two) break;
        const int gg = g + h * bs;
        const long long i = kb + 2 * gg;
        // d, p and q are not read again before the next iteration has
        // streamed a whole working set (larger than the last-level cache):
        // non-temporal, so that they do not evict z, the dictionaries and
        // the partials
        dg[h] = *reinterpret_cast<const T2 *>(dinv + 2 * gg);
        di[h].x = __builtin_nontemporal_load(dv + i);
        di[h].y = __builtin_nontemporal_load(dv + i + 1);
        zi[h] = *reinterpret_cast<const T2 *>(zv + i);
        pi[h].x = __builtin_nontemporal_load(p + i);
        pi[h].y = __builtin_nontemporal_load(p + i + 1);
        qi[h].x = __builtin_nontemporal_load(q + i);
        qi[h].y = __builtin_nontemporal_load(q + i + 1);
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (h == 1 && !two) break;
        const long long i = kb + 2 * (g + h * bs);
        di[h].x = fma(a, pi[h].x, di[h].x);
        di[h].y = fma(a, pi[h].y, di[h].y);
        zi[h].x = fma(-a, dg[h].x * qi[h].x, zi[h].x);
        zi[h].y = fma(-a, dg[h].y * qi[h].y, zi[h].y);
        __builtin_nontemporal_store(di[h].x, dv + i);
        __builtin_nontemporal_store(di[h].y, dv + i + 1);
        *reinterpret_cast<T2 *>(zv + i) = zi[h];
        const TS rx = zi[h].x / dg[h].x, ry = zi[h].y / dg[h].y;  // r = z / dinv
        vu[k] = fma((double)rx, (double)rx, vu[k]);
        vu[k] = fma((double)ry, (double)ry, vu[k]);
        vu[K + k] = fma((double)rx, (double)zi[h].x, vu[K + k]);
        vu[K + k] = fma((double)ry, (double)zi[h].y, vu[K + k]);
      }
    }
    if ((rows & 1) && tid == 0) {
      const long long i = kb + rows - 1;
      const TS dg = dinv[rows - 1];
      const TS di = fma(a, p[i], dv[i]);
      const TS zi = fma(-a, dg * q[i], zv[i]);
      dv[i] = di;
      zv[i] = zi;
      const TS ri = zi / dg;
      vu[k] = fma((double)ri, (double)ri, vu[k]);
      vu[K + k] = fma((double)ri, (double)zi, vu[K + k]);
    }
  }
  block_sum<2 * K>(vu, lds);
  if (tid == 0) {
    const SlabPart part = slab_part(P, slab);
    for (int i = 0; i < 2 * K; i++) part[K + i] = vu[i];
  }
}

// ---------------------------------------------------------------------------
// Kernel R (reads st[1], writes st[2]).
// PH_RES: the FP64 residual r = -K x64 over the FP64 dictionary rows (the
// assembled matrix; the Newton / refinement residual of
// micro/src/assembly.c:33-82; boundary rows keep a zero residual, :59-74),
// and the start of the next PCG sweep (d = 0, z = dinv r, p = q = 0): every
// refinement sweep solves the correction equation from zero.  The relative
// target refers to the cold-start right-hand side (interior displacement
// zero, the ||b|| of the first solve); an RVE started from the affine field
// or warm-started measures it at sweep 0 from the boundary neighbours alone.
// PH_AVG: per-slab partial volume averages of strain and stress in FP64,
// sum_gp (.) w_gp per element through sum_gp B w, one load case at a time.
template <typename TS, int K>
__device__ void res_slab(const EngPtrs &P, int slab, int slot, const RveDev *R, bool need_cold,
                         double (&lds)[RedLds<K>::size])
{
  const int tid = threadIdx.x, RS = 3 * P.SN;
  const int n = P.slab_node0[slab] + tid;
  const bool valid = n < R->nrow / 3;
  double v[4 * K];
#pragma unroll
  for (int i = 0; i < 4 * K; i++) v[i] = 0.0;
  if (valid) {
    const int nx = R->nx, ny = R->ny, nz = R->nz, nxy = nx * ny;
    const int ix = n % nx, iy = (n / nx) % ny, iz = n / nxy;
    const int cls = node_class(ix, iy, iz, nx, ny, nz);
    // K x64 over the node's 27 FP64 dictionary blocks (the assembled FP64
    // rows); boundary rows keep a zero residual
    double r64[3][K], rc64[3][K];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int k = 0; k < K; k++) {
        r64[d][k] = 0.0;
        rc64[d][k] = 0.0;
      }
    if (cls == 13) {
      const double *D = P.dict64 + ((size_t)slot * kDictPat + P.npat[(long long)slab * P.SN + tid]) *
                                       kDictEntry;
#pragma unroll 1
      for (int key = 0; key < 27; key++) {
        const int dx = key % 3 - 1, dy = (key / 3) % 3 - 1, dz = key / 9 - 1;
        const int m = n + dx + dy * nx + dz * nxy;
        const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
        const bool nb_bc = jx == 0 || jx == nx - 1 || jy == 0 || jy == ny - 1 || jz == 0 || jz == nz - 1;
        double vb[3][4];
#pragma unroll
        for (int dd = 0; dd < 3; dd++) load_vec<double, 4>(D + key * kDictBlk + dd * 4, vb[dd]);
        double xv[3][K];
        gather_node<double, K>(P.x64, node_vec(P, R->first_slab, m, K), RS, xv);
#pragma unroll
        for (int dd = 0; dd < 3; dd++)
#pragma unroll
          for (int d = 0; d < 3; d++)
#pragma unroll
            for (int k = 0; k < K; k++) r64[d][k] = fma(vb[dd][d], xv[dd][k], r64[d][k]);
        if (need_cold && nb_bc)
#pragma unroll
          for (int dd = 0; dd < 3; dd++)
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
              for (int k = 0; k < K; k++) rc64[d][k] = fma(vb[dd][d], xv[dd][k], rc64[d][k]);
      }
    }
    const long long nb = node_vec(P, R->first_slab, n, K);
    const TS *dinv = (const TS *)P.dinv + (long long)slab * RS + 3 * tid;
    TS ri[3][K], zi[3][K], zero[3][K];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const TS dg = dinv[d];
#pragma unroll
      for (int k = 0; k < K; k++) {
        const double rk = -r64[d][k];
        ri[d][k] = (TS)rk;
        zi[d][k] = dg * ri[d][k];
        zero[d][k] = (TS)0;
        v[k] = fma(rk, rk, v[k]);
        v[K + k] = fma((double)ri[d][k], (double)ri[d][k], v[K + k]);
        v[2 * K + k] = fma((double)ri[d][k], (double)zi[d][k], v[2 * K + k]);
        v[3 * K + k] = fma(rc64[d][k], rc64[d][k], v[3 * K + k]);
      }
    }

    scatter_node<TS, K>((TS *)P.z, nb, RS, zi);
    scatter_node<TS, K>((TS *)P.d, nb, RS, zero);
    scatter_node<TS, K>((TS *)P.p, nb, RS, zero);
    scatter_node<TS, K>((TS *)P.q, nb, RS, zero);
  }
  block_sum<4 * K>(v, lds);
  if (tid == 0) {
    const SlabPart part = slab_part(P, slab);
    for (int i = 0; i < 4 * K; i++) part[3 * K + i] = v[i];
  }
}

template <typename TS, int K>
__device__ void avg_slab(const EngPtrs &P, int slab, int slot, const RveDev *R,
                         double (&lds)[RedLds<K>::size])
{
  __shared__ double bw[144];
  __shared__ double cm[kNmat * 36];
  const int tid = threadIdx.x;
  const int nslab = R->nslab;
  const SlabPart part = slab_part(P, slab);
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
          const double xv = P.x64[node_vec(P, R->first_slab, node, K) + (long long)k * 3 * P.SN + dd];
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
}

// Kernel R's work on one slab of an RVE being solved.
template <typename TS, int K>
__device__ void r_slab(const EngPtrs &P, int slab, double (&lds)[RedLds<K>::size])
{
  const int tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  RveDev *R = P.rves + slot;
  const SolveState *S = &R->st[1];
  const int phase = S->phase;
  if (slab == R->first_slab && tid == 0) {
    SolveState N = *S;
    if (phase == PH_RES) N.phase = PH_RESRED;
    else if (phase == PH_AVG) N.phase = PH_AVGRED;
    R->st[2] = N;
  }
  if (phase == PH_RES) res_slab<TS, K>(P, slab, slot, R, (R->warm || R->affine) && S->sweep == 0, lds);
  else if (phase == PH_AVG) avg_slab<TS, K>(P, slab, slot, R, lds);
  else if (phase == PH_INNER && R->big && slab == R->first_slab) {
    // B's r.r / r.z totals for kernel A (group_sum_cols over SN lanes, as A
    // would form them)
    double s2[2 * K];
    group_sum_cols<2 * K>(slab_part(P, R->first_slab).p, R->nslab, P.part_stride, K, s2, lds);
    if (tid == 0)
      for (int i = 0; i < 2 * K; i++) R->rr_rz[i] = s2[i];
  }
}

// Runs work(slab) for the slabs of the active list for which want(slab, R)
// holds, on a grid of at most kRGrid workgroups (kernels R and S, which have
// work on few slabs in most launches, so a workgroup per slab would mostly
// start and exit).  Workgroup b takes the positions j = b / 8 + m * (grid / 8)
// of its XCD's part (b % 8) of the active list (the split of active_in).  Per
// round its lanes look up one position each and collect the slabs with work
// in LDS; the workgroup then works on those, in any order (each writes only
// its own slab's outputs).
template <typename Want, typename Work>
__device__ void for_slabs_with_work(const EngPtrs &P, Want want, Work work)
{
  __shared__ int slabs[256];
  __shared__ int nwork;
  const int *list = P.act_list;
  const int n = list[0];
  const int per = (n + kXcd - 1) / kXcd;
  const int g8 = gridDim.x / kXcd, part = blockIdx.x % kXcd, j0 = blockIdx.x / kXcd;
  const int mine = (j0 < per) ? (per - 1 - j0) / g8 + 1 : 0;
  const int tid = threadIdx.x;
  for (int m0 = 0; m0 < mine; m0 += blockDim.x) {
    if (tid == 0) nwork = 0;
    __syncthreads();
    const int m = m0 + tid;
    const int i = part * per + j0 + m * g8;
    if (m < mine && i < n) {
      const int slab = list[1 + i];
      const int slot = P.slab_rve[slab];
      if (slot >= 0 && want(slab, P.rves + slot)) slabs[atomicAdd(&nwork, 1)] = slab;
    }
    __syncthreads();
    const int cnt = nwork;
    for (int w = 0; w < cnt; w++) work(slabs[w]);
    __syncthreads();  // slabs and nwork are rewritten by the next round
  }
}

// Kernel R: an RVE's first slab passes the state on (and, for big RVEs in a
// sweep, sums r.r / r.z); the residual and averaging phases work on every
// slab.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_r(EngPtrs P, SolveParams sp)
{
  __shared__ double lds[RedLds<K>::size];
  for_slabs_with_work(
      P,
      [](int slab, const RveDev *R) {
        const int ph = R->st[1].phase;
        return slab == R->first_slab || ph == PH_RES || ph == PH_AVG;
      },
      [&](int slab) { r_slab<TS, K>(P, slab, lds); });
}

// Kernel S, launched after A only when the launch's group holds a big RVE
// (RveDev::big): sums A's p.q partials of each big RVE in a sweep once into
// RveDev::pq (group_sum_cols over SN lanes, as every workgroup of B would),
// so that kernel B reads K totals instead of re-summing nslab partials.
template <typename TS, int K>
__global__ __launch_bounds__(256) void k_step_s(EngPtrs P)
{
  __shared__ double lds[16 * K];
  for_slabs_with_work(
      P,
      [](int slab, const RveDev *R) {
        return R->big && slab == R->first_slab && R->st[0].phase == PH_INNER;
      },
      [&](int slab) {
        RveDev *R = P.rves + P.slab_rve[slab];
        double s1[K];
        group_sum_cols<K>(slab_part(P, R->first_slab).p, R->nslab, P.part_stride, 0, s1, lds);
        if (threadIdx.x == 0)
          for (int k = 0; k < K; k++) R->pq[k] = s1[k];
      });
}

// Stand-alone copy of the INNER-phase SpMV (w = A z) over every resident RVE,
// written to q; used only when no RVE is being solved.
template <typename TS, int K, bool DICT>
__global__ __launch_bounds__(256) void k_spmv_time(EngPtrs P)
{
  const int slab = xcd_slab(), tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  const RveDev *R = P.rves + slot;
  const int n = P.slab_node0[slab] + tid;
  const DictCache<TS> *dc = dict_cache<TS, DICT>(P, R, slot, slab);
  if (n >= R->nrow / 3) return;
  TS w[3][K], zc[3][K];
  slab_node_spmv<TS, K, DICT>(P, R, slot, slab, tid, n, dc, w, zc);
  scatter_node<TS, K>((TS *)P.q, node_vec(P, R->first_slab, n, K), 3 * P.SN, w);
}

// Stand-alone replica of kernel A's INNER work (inner_a, beta = 1/2 for every
// load case) over every resident RVE; overwrites p, q and the partials, so it
// is used only when no RVE is being solved.
template <typename TS, int K, bool DICT>
__global__ __launch_bounds__(256) void k_inner_time(EngPtrs P)
{
  __shared__ double lds[RedLds<K>::size];
  __shared__ double beta[K];
  const int slab = xcd_slab(), tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  if (P.a_lds > 0 && P.rves[slot].mcpp > 0) return;  // k_march_time's
  if (tid < K) beta[tid] = 0.5;
  __syncthreads();
  inner_a<TS, K, DICT>(P, P.rves + slot, slot, slab, 0u, beta, lds);
}

// The same replica for the plane march (march_tile over the tiles of every
// resident RVE with mcpp > 0); the launch's dynamic LDS covers them all.
template <typename TS, int K, bool DICT>
__global__ __launch_bounds__(256) void k_march_time(EngPtrs P)
{
  __shared__ double lds[MarchLds<K>::size];
  __shared__ double beta[K];
  const int slab = xcd_slab(), tid = threadIdx.x;
  const int slot = P.slab_rve[slab];
  if (slot < 0) return;
  const RveDev *R = P.rves + slot;
  const int t = slab - R->first_slab;
  if (R->mcpp == 0 || t >= R->mnt * R->mncg) return;  // slab by slab: k_inner_time
  if (tid < K) beta[tid] = 0.5;
  __syncthreads();
  march_tile<TS, K, DICT>(P, R, slot, t, 0u, beta, lds);
}

// ---------------------------------------------------------------------------
// launchers (explicit instantiations for the supported load-case counts)

// nkern launches of the cycle A (0), B (1), R (2), starting at `first`.
template <typename TS, int K, bool DICT>
static int launch_steps(const EngPtrs &P, const SolveParams &sp, int grid, int first, int nkern, hipStream_t s)
{
  if (grid % kXcd || P.a_grid % kXcd) return MM_ERR_ARG;  // the XCD-aware order needs whole groups of 8
  if (P.a_lds < 0 || P.a_lds > kMarchMaxLds) return MM_ERR_ARG;
  if (first < 0 || first > 2 || nkern < 0) return MM_ERR_ARG;
  for (int i = first; i < first + nkern; i++) {
    switch (i % 3) {
      case 0:
        // with the march on (a_lds > 0), its tiles and, if any RVE is
        // processed slab by slab (a_slab), k_step_a for those
        if (P.a_lds > 0 && P.a_grid > 0) {
          if (DICT) k_step_a_march_dict<TS, K><<<P.a_grid, P.SN, P.a_lds, s>>>(P, sp);
          else k_step_a_march_full<TS, K><<<P.a_grid, P.SN, P.a_lds, s>>>(P, sp);
        }
        if (P.a_lds == 0 || P.a_slab) k_step_a<TS, K, DICT><<<grid, P.SN, 0, s>>>(P, sp);
        if (P.s_big) k_step_s<TS, K><<<grid < kRGrid ? grid : kRGrid, P.SN, 0, s>>>(P);
        break;
      case 1: k_step_b<TS, K><<<grid, P.SN, 0, s>>>(P, sp); break;
      default: k_step_r<TS, K><<<grid < kRGrid ? grid : kRGrid, P.SN, 0, s>>>(P, sp); break;
    }
  }
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename TS, int K>
static int launch_assemble(const EngPtrs &P, int nblk, const int *slab_list, int mode, hipStream_t s)
{
  if (nblk <= 0) return 0;
  k_assemble<TS, K><<<nblk, P.SN, 0, s>>>(nblk, slab_list, P.slab_rve, P.slab_node0, P.rves, P.ke_tab,
                                          P.elem_mat, const_cast<int *>(P.bcols), (TS *)P.vals,
                                          (TS *)P.dinv, const_cast<uint8_t *>(P.npat), P.x64, P.SN,
                                          mode);
  MM_HIP(hipGetLastError());
  return 0;
}

template <typename TS, int K, bool DICT>
static int launch_time(const EngPtrs &P, int grid, int what, hipStream_t s)
{
  if (what == 1) {
    if (P.a_lds > 0) k_march_time<TS, K, DICT><<<grid, P.SN, P.a_lds, s>>>(P);
    if (P.a_lds == 0 || P.a_slab) k_inner_time<TS, K, DICT><<<grid, P.SN, 0, s>>>(P);
  } else {
    k_spmv_time<TS, K, DICT><<<grid, P.SN, 0, s>>>(P);
  }
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



int eng_launch_steps(bool fp64, int K, bool dict, const EngPtrs &P, const SolveParams &sp, int grid,
                     int first, int nkern, hipStream_t s)
{
  if (dict) {
#define CALL_STEPS(T, KK) launch_steps<T, KK, true>(P, sp, grid, first, nkern, s)
    MM_DISPATCH(fp64, K, CALL_STEPS);
#undef CALL_STEPS
  }
#define CALL_STEPS(T, KK) launch_steps<T, KK, false>(P, sp, grid, first, nkern, s)
  MM_DISPATCH(fp64, K, CALL_STEPS);
#undef CALL_STEPS
}

int eng_launch_dict(bool fp64, int n, const int *slots, const double *ke_tab, void *dict, hipStream_t s)
{
  if (n <= 0) return 0;
  if (fp64) k_dict<double><<<n * kDictPat, 384, 0, s>>>(n, slots, ke_tab, (double *)dict);
  else k_dict<float><<<n * kDictPat, 384, 0, s>>>(n, slots, ke_tab, (float *)dict);
  MM_HIP(hipGetLastError());
  return 0;
}

int eng_launch_assemble(bool fp64, int K, const EngPtrs &P, int nblk, const int *slab_list, int mode,
                        hipStream_t s)
{
#define CALL_ASM(T, KK) launch_assemble<T, KK>(P, nblk, slab_list, mode, s)
  MM_DISPATCH(fp64, K, CALL_ASM);
#undef CALL_ASM
}

int eng_launch_time(bool fp64, int K, bool dict, const EngPtrs &P, int grid, int what, hipStream_t s)
{
  if (dict) {
#define CALL_TIME(T, KK) launch_time<T, KK, true>(P, grid, what, s)
    MM_DISPATCH(fp64, K, CALL_TIME);
#undef CALL_TIME
  }
#define CALL_TIME(T, KK) launch_time<T, KK, false>(P, grid, what, s)
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
