// mm_internal.h -- internal declarations shared by the HIP translation units
// of libmadmax_amd.so.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "micro_api.h"

#define MM_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "madmax_amd: HIP error %s (%d) at %s:%d: %s\n", hipGetErrorName(e_), \
              (int)e_, __FILE__, __LINE__, #call);                                    \
      return MM_ERR_HIP;                                                              \
    }                                                                                 \
  } while (0)

namespace mm {

constexpr int kW = 81;       // ELL width of the 3D hex8 Jacobian (micro/src/alloc.c:6)
constexpr int kNd = 24;      // element dofs, 8 nodes x 3
constexpr int kKe = kNd * kNd;
constexpr int kNmat = 2;     // MATRIX, FIBER

// ---------------------------------------------------------------------------
// Boundary classes of a structured hex8 node.  Along each axis a node is at
// the low face (0), inside (1) or at the high face (2); class = cx + 3cy + 9cz.
// For every class the ELL column order of a row is the order in which the
// reference's ell_add_vals first touches each neighbour node (elements in
// ascending id, element-local node order, src/ell.c:46-84), so it is a fixed
// list of neighbour offsets.  Each slot also lists the elements that add into
// it, in ascending element id, with the local node indices a (row node) and b
// (column node) into that element's Ke.
struct SlotTables {
  int nslot[27];                 // neighbour nodes of the class (8..27)
  uint8_t off[27][27];           // packed (dx+1) + 3(dy+1) + 9(dz+1)
  uint8_t ncontrib[27][27];      // elements adding into the slot (1..8)
  uint16_t contrib[27][27][8];   // (ox+1) | (oy+1)<<1 | (oz+1)<<2 | a<<3 | b<<6
  uint8_t diag_slot[27];         // slot of the node itself
  uint8_t nelem[27];             // elements around the node (1..8)
  uint8_t elem[27][8];           // (ox+1) | (oy+1)<<1 | (oz+1)<<2 | a<<3, ascending id
  int8_t key_slot[27][27];       // slot of neighbour key (dx+1)+3(dy+1)+9(dz+1), -1 if none
};

void build_slot_tables(SlotTables *t);

// element-local node index from local coordinates (src/fem.c:324-361 order)
__host__ __device__ inline int local_node(int lx, int ly, int lz)
{
  const int tab[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // index lx + 2ly + 4lz
  return tab[lx + 2 * ly + 4 * lz];
}
__host__ __device__ inline void local_coords(int a, int *lx, int *ly, int *lz)
{
  const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  const int cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  *lx = cx[a];
  *ly = cy[a];
  *lz = a >> 2;
}

// ---------------------------------------------------------------------------
// Solver phases of one resident RVE.  One solver step is three launches,
// A -> B -> R, over every resident slab:
//   A  PH_RESRED: reduce the residual partials, start a PCG sweep or go to
//                 PH_AVG; then as PH_INNER
//      PH_INNER:  reduce r.r and r.z of the last B; converged -> x += d and
//                 PH_RES; else SpMV, p, q and partial p.q
//      PH_AVGRED: reduce the averaging partials, write the result, PH_DONE
//   B  PH_INNER:  reduce p.q, alpha; d, z (z = dinv r; r is not stored) and
//                 partial r.r, r.z
//   R  PH_RES:    FP64 residual, partial norms -> PH_RESRED
//      PH_AVG:    partial volume averages -> PH_AVGRED
// Every reduction is done by each workgroup of the consuming launch over the
// producing launch's per-slab partials in slab order, so all workgroups of an
// RVE agree bit for bit and no workgroup waits on, or fences for, another.
enum Phase : int {
  PH_FREE = 0,
  PH_RES = 1,
  PH_RESRED = 2,
  PH_INNER = 3,
  PH_AVG = 4,
  PH_AVGRED = 5,
  PH_DONE = 6
};

// Per-RVE solver scalars.  RveDev holds three copies: st[0] is written by A,
// st[1] by B, st[2] by R (and by admission); each launch reads only its
// predecessor's copy, and only the workgroup of the RVE's first slab writes.
struct SolveState {
  int phase, sweep, its, its_total;
  int status;
  unsigned done_in, done_out;
  int pad_;
  double beta[MM_MAX_RHS], rz[MM_MAX_RHS];
  double bnorm_in[MM_MAX_RHS];  // absolute PCG target of the current sweep
  double r0[MM_MAX_RHS], rnorm[MM_MAX_RHS];
};

// Admission-time description of a resident RVE plus its solver state.
struct RveDev {
  // geometry and materials (written at admission)
  int nex, ney, nez, nx, ny, nz;
  int nrow, nelm, nslab, first_slab;
  long long row_base;   // global row of local row 0 (first_slab * 3 * SN)
  long long elem_base;  // offset into the element-material arena
  double hx, hy, hz, vol;
  int nfib, tangent_mode, nload, warm;
  int affine;  // cold start from the affine field u = E X (else interior zero)
  // plane-march tiling of kernel A: the RVE's nodes as mnt = mcpp * mnzc
  // blocks, mcpp in-plane chunks of at most SN nodes times mnzc runs of at
  // most mL z-planes (even splits), each marched once per group of load
  // cases (mncg groups of march_cases): tile t = block + mnt * group, t <
  // mnt * mncg <= nslab, is processed by the workgroup of slab
  // first_slab + t.  mcpp == 0: kernel A works slab by slab.
  int mcpp, mL, mnzc, mnt, mncg;
  // big != 0 (nslab >= mm_engine_config::big_slabs): during PCG sweeps, kernel R sums B's r.r /
  // r.z partials once into rr_rz and kernel S sums A's p.q partials once
  // into pq (in the RVE's first slab's workgroup, in the order every
  // workgroup of the consuming kernel would use), and kernels A and B read
  // the totals instead of each workgroup re-summing nslab partials.
  int big;
  double rr_rz[2 * MM_MAX_RHS];
  double pq[MM_MAX_RHS];
  double fx[MM_MAX_FIBERS], fy[MM_MAX_FIBERS], fr[MM_MAX_FIBERS];
  double young[kNmat], poisson[kNmat];
  double load[MM_MAX_RHS][MM_NVOI];
  // warm start (warm != 0): interior displacement of load k starts at
  // sum_j wcoef[k][j] x_old[j] + E(dload[k]) X, x_old the kept displacements
  double wcoef[MM_MAX_RHS][MM_MAX_RHS];
  double dload[MM_MAX_RHS][MM_NVOI];
  SolveState st[3];
};

// Results written by the device into host-pinned memory.
struct RveOut {
  int phase;   // mirrors RveDev::phase once the RVE is finished
  int status, its_total, sweeps;
  double residual_rel;
  double stress[MM_MAX_RHS][MM_NVOI];
  double strain[MM_MAX_RHS][MM_NVOI];
  double c[MM_NVOI * MM_NVOI];
};

struct SolveParams {
  int nrhs, slab_nodes;
  double rtol, abs_tol, inner_rtol;
  int max_inner_its, max_sweeps;
};

// Device pointers of the engine arena handed to every solver kernel.
struct EngPtrs {
  const int *bcols;  // [slab][27][SN] neighbour node of each slot block
  const void *vals;  // [slab][81][3*SN]
  const void *dinv;
  double *x64;
  void *d, *z, *p, *q;  // residual kept only as z = dinv r
  double *part;           // [value][slab] partial sums, value stride part_stride
  long long part_stride;  // slabs of the arena
  const int *slab_rve, *slab_node0;
  RveDev *rves;
  RveOut *out;
  const double *ke_tab, *c_tab, *bw_tab;
  const uint8_t *elem_mat;
  const uint8_t *npat;  // [slab][SN] material pattern of the node's 8 elements
  const int *act_list;  // [0]: number of slabs being solved, [1..]: those slabs
  const void *dict;     // [rve slot][256][27][12] interior 3-row ELL blocks (TS)
  const double *dict64; // the same in FP64 (the residual's matrix)
  const int *act_tile;  // [0]: tiles of kernel A's plane march, [1..]: their slabs
  int SN;        // nodes per slab = lanes per workgroup (a power of two)
  int sn_shift;  // log2(SN)
  int a_grid;    // grid of the plane march over act_tile (a_lds > 0)
  int a_lds;     // dynamic LDS bytes of the plane march; 0: march off
  int a_slab;    // march on: some RVEs (mcpp == 0) need k_step_a as well
  int s_big;     // the group holds an RVE with big != 0: kernel S runs after A
};


// Both matrix storages keep a node's 27 neighbour blocks in neighbour-key
// order, key = (dx+1) + 3(dy+1) + 9(dz+1) (z-major, x fastest), so the SpMV
// visits the three x-neighbours of a (dy, dz) strip one after the other.
// key_slot maps a key to the reference's slot of that block in the node's
// row (first-touch order), which fixes the values and the ELL download.
// Dictionary ELL: entry (slot, pattern) holds, for the 27 neighbour keys of
// an interior node, v[d'][d] (value of row 3n+d, column 3m+d') at
// [key * 12 + d' * 4 + d] (padded to 16-byte rows).
constexpr int kDictPat = 256;
constexpr int kDictBlk = 12;
constexpr int kDictEntry = 27 * kDictBlk;

// Load cases staged in LDS per round of the plane march.  Two keep an FP32
// march of a 30^3 RVE (window 318 nodes) with its dictionary cache under
// 40 KB, four workgroups per CU; three cases per round fit only three.
constexpr int march_cases(int tsz) { return tsz == 8 ? 1 : 2; }
// Default of mm_engine_config::big_slabs: RVEs of at least this many slabs
// get their dot-product totals from kernels R and S (RveDev::big); below it
// the re-summing in every workgroup costs less than the extra latency of
// one workgroup in R and of kernel S's launch.
constexpr int kBigSlabs = 1024;
// Dictionary entries a march tile keeps in LDS, and the bytes of that cache
// (entries, then one slot byte per pattern), a multiple of 16.
constexpr int kMarchDictCap = 10;
constexpr int march_dict_bytes(int tsz) { return (kMarchDictCap * kDictEntry * tsz + kDictPat + 15) / 16 * 16; }
// LDS of the plane march: the dictionary cache (dictionary storage), then a
// ring of three planes of the window (the largest in-plane chunk of the even
// split, plus nx + 1 nodes either side) for the round's load cases, three
// rows each.
inline int march_lds_bytes(int SN, int nx, int ny, int K, int tsz, bool dict)
{
  const int g = K < march_cases(tsz) ? K : march_cases(tsz);
  const long long nxy = (long long)nx * ny, cpp = (nxy + SN - 1) / SN;
  const int chunk = (int)((nxy + cpp - 1) / cpp);  // the largest chunk of the even split
  return (dict ? march_dict_bytes(tsz) : 0) + 3 * g * 3 * (chunk + 2 * nx + 2) * tsz;
}
constexpr int kMarchMaxLds = 64 * 1024;

// assembly modes
enum AsmMode : int { ASM_FULL = 0, ASM_X_ONLY = 1, ASM_DICT = 2 };

// launchers (engine_kernels.hip)
int upload_slot_tables(hipStream_t s);
bool eng_supported_rhs(int K);
// nkern solver launches of the cycle A (0) -> B (1) -> R (2), from `first`
int eng_launch_steps(bool fp64, int K, bool dict, const EngPtrs &P, const SolveParams &sp, int grid,
                     int first, int nkern, hipStream_t s);
int eng_launch_dict(bool fp64, int n, const int *slots, const double *ke_tab, void *dict, hipStream_t s);
int eng_launch_assemble(bool fp64, int K, const EngPtrs &P, int nblk, const int *slab_list, int mode,
                        hipStream_t s);
// what: 0 = the SpMV w = A z alone (k_spmv_time), 1 = kernel A's INNER work
int eng_launch_time(bool fp64, int K, bool dict, const EngPtrs &P, int grid, int what, hipStream_t s);
int eng_launch_admit(int n, const int *slots, const RveDev *states, RveDev *rves, int *slab_rve,
                     int *slab_node0, int SN, hipStream_t s);
int eng_launch_ke(int n, const int *work, const RveDev *rves, double *ke_tab, double *c_tab_out,
                  double *bw_tab, hipStream_t s);
int eng_launch_phase(int nblk, const int2 *work, const RveDev *rves, uint8_t *elem_mat, hipStream_t s);
constexpr int kPartDoubles = 96;


}  // namespace mm
