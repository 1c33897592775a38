// fe.hip -- per-element FE assembly on the GPU: element stiffness matrices
// and residuals from a tangent / stress per gauss point, summed into the
// reference's ELL (and residual vector) by one lane per node in ascending
// element order.  Serves the micro's request path with re-assembly per
// Newton iteration (micro/src/homogenize.c:175-225), the periodic boundary
// condition (micro/src/assembly.c:175-220, 272-354) and the macro mesh
// (macro/src/assembly.c:5-222).  Every kernel is FP64 with multiply-add
// contraction off, in the reference's operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fe.h"
#include "mm_device.h"

namespace mm {

namespace {

constexpr int kNnz = 81;  // ELL width of the 3D hex8 Jacobian (micro/src/alloc.c:6)
constexpr int kEB = 8;    // elements per wavefront in the stiffness kernel
constexpr double kGauss = 0.577350269189626;  // src/fem.c:324

// Device view of the mesh.
struct FeDev {
  int nex, ney, nez, nx, ny, nz;
  int nelm, nunk, nrow, periodic, iso, flush;
  double hx, hy, hz, vol;
  const double *coords;      // [structured node][3] (iso)
  const uint8_t *elem_mat;   // [elem]
  const double *cmat;        // [mat][36]
  const double *c_gp;        // [elem][8][36] or null
  const uint8_t *dir;        // [dof] Dirichlet flag
  int zero_cols;
};

// hex8 master derivative ds[a][i] at gauss point g (src/fem.c:364-405 order:
// node a and gauss point g at signs (-,-,-) (+,-,-) (+,+,-) (-,+,-), then z +)
__device__ __forceinline__ double hex_sign(int a, int d)
{
  if (d == 0) return (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0;
  if (d == 1) return (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0;
  return (a >= 4) ? 1.0 : -1.0;
}
__device__ __forceinline__ void master_ds(int a, int g, double (&ds)[3])
{
#pragma clang fp contract(off)
  double f[3];
  for (int d = 0; d < 3; d++) f[d] = 1.0 + hex_sign(a, d) * (hex_sign(g, d) * kGauss);
  ds[0] = hex_sign(a, 0) * f[1] * f[2] * 0.125;
  ds[1] = hex_sign(a, 1) * f[0] * f[2] * 0.125;
  ds[2] = hex_sign(a, 2) * f[0] * f[1] * 0.125;
}

// Structured node of local node a of element e.
__device__ __forceinline__ void elem_node(const FeDev &F, int e, int a, int &ix, int &iy, int &iz)
{
  int lx, ly, lz;
  local_coords(a, &lx, &ly, &lz);
  ix = e % F.nex + lx;
  iy = (e / F.nex) % F.ney + ly;
  iz = e / (F.nex * F.ney) + lz;
}
__device__ __forceinline__ int unknown_of(const FeDev &F, int ix, int iy, int iz)
{
  if (F.periodic) return (ix % F.nex) + (iy % F.ney) * F.nex + (iz % F.nez) * F.nex * F.ney;
  return ix + iy * F.nx + iz * F.nx * F.ny;
}
__device__ __forceinline__ void node_coord(const FeDev &F, int ix, int iy, int iz, double (&X)[3])
{
  if (F.iso) {
    const double *c = F.coords + 3 * ((long long)ix + (long long)iy * F.nx + (long long)iz * F.nx * F.ny);
    X[0] = c[0];
    X[1] = c[1];
    X[2] = c[2];
  } else {
    X[0] = ix * F.hx;
    X[1] = iy * F.hy;
    X[2] = iz * F.hz;
  }
}

// Physical derivatives dsh[a][i] and weight at gauss point g of element e:
// structured (ds * 2 / h, w = vol_e / 8; src/fem.c:42-54 extended to 3D) or
// isoparametric (J = sum_n ds X, cofactor inverse, dsh = J^-1 ds, w = |det J|
// times the unit hex8 weight; src/fem.c:411-471, macro/src/assembly.c:115).
__device__ void gp_dsh(const FeDev &F, int e, int g, double (&dsh)[8][3], double &w)
{
#pragma clang fp contract(off)
  if (!F.iso) {
    const double h[3] = {F.hx, F.hy, F.hz};
    for (int a = 0; a < 8; a++) {
      double ds[3];
      master_ds(a, g, ds);
      for (int d = 0; d < 3; d++) dsh[a][d] = ds[d] * 2 / h[d];
    }
    const double vol = h[0] * h[1] * h[2];
    w = vol / 8;
    return;
  }
  double X[8][3], ds[8][3];
  for (int a = 0; a < 8; a++) {
    int ix, iy, iz;
    elem_node(F, e, a, ix, iy, iz);
    node_coord(F, ix, iy, iz, X[a]);
    master_ds(a, g, ds[a]);
  }
  double J[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int n = 0; n < 8; n++) s += ds[n][i] * X[n][j];
      J[i][j] = s;
    }
  const double c00 = +J[1][1] * J[2][2] - J[2][1] * J[1][2], c01 = -J[1][0] * J[2][2] + J[2][0] * J[1][2],
               c02 = +J[1][0] * J[2][1] - J[2][0] * J[1][1];
  const double c10 = -J[0][1] * J[2][2] + J[2][1] * J[0][2], c11 = +J[0][0] * J[2][2] - J[2][0] * J[0][2],
               c12 = -J[0][0] * J[2][1] + J[2][0] * J[0][1];
  const double c20 = +J[0][1] * J[1][2] - J[1][1] * J[0][2], c21 = -J[0][0] * J[1][2] + J[1][0] * J[0][2],
               c22 = +J[0][0] * J[1][1] - J[1][0] * J[0][1];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double iJ[3][3] = {{c00 / det, c10 / det, c20 / det}, {c01 / det, c11 / det, c21 / det},
                           {c02 / det, c12 / det, c22 / det}};
  for (int a = 0; a < 8; a++)
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 3; j++) s += iJ[i][j] * ds[a][j];
      dsh[a][i] = s;
    }
  w = fabs(det) * 1.0;
}

// B[v][3a + d] from dsh[a][.] (Voigt xx, yy, zz, xy, yz, xz; engineering
// shears; micro/src/init.c:40-50 extended to 3D)
__device__ __forceinline__ double bval(const double (&dsh)[3], int v, int d)
{
  switch (v) {
    case 0: return d == 0 ? dsh[0] : 0.0;
    case 1: return d == 1 ? dsh[1] : 0.0;
    case 2: return d == 2 ? dsh[2] : 0.0;
    case 3: return d == 0 ? dsh[1] : (d == 1 ? dsh[0] : 0.0);
    case 4: return d == 1 ? dsh[2] : (d == 2 ? dsh[1] : 0.0);
    default: return d == 0 ? dsh[2] : (d == 2 ? dsh[0] : 0.0);
  }
}

// Element stiffness k[i][j] = sum_gp sum_k sum_h B[h][i] C[h][k] B[k][j] w
// (micro/src/assembly.c:97-107, macro/src/assembly.c:173-181), one wavefront
// per block of kEB elements: lane 8 el + g computes the derivatives of its
// element's gauss point g into LDS, then the lanes share the 576 entries of
// each element.  Terms with a zero factor are skipped (adding an exact zero
// leaves every sum unchanged).
__global__ __launch_bounds__(64) void k_fe_ke(FeDev F, double *__restrict__ ke_out)
{
#pragma clang fp contract(off)
  __shared__ double s_dsh[kEB][8][8][3];
  __shared__ double s_w[kEB][8];
  __shared__ double s_c[kEB][8][36];
  const int e0 = blockIdx.x * kEB, l = threadIdx.x;
  {
    const int el = l / 8, g = l % 8, e = e0 + el;
    if (e < F.nelm) {
      double dsh[8][3], w;
      gp_dsh(F, e, g, dsh, w);
      for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++) s_dsh[el][g][a][d] = dsh[a][d];
      s_w[el][g] = w;
      const double *c = F.c_gp ? F.c_gp + ((long long)e * 8 + g) * 36 : F.cmat + (size_t)F.elem_mat[e] * 36;
      for (int i = 0; i < 36; i++) s_c[el][g][i] = c[i];
    }
  }
  __syncthreads();
  for (int idx = l; idx < kEB * 576; idx += 64) {
    const int el = idx / 576, ent = idx % 576, e = e0 + el;
    if (e >= F.nelm) break;
    const int i = ent / 24, j = ent % 24, a = i / 3, d = i % 3, b = j / 3, dd = j % 3;
    double acc = 0.0;
    for (int g = 0; g < 8; g++) {
      const double w = s_w[el][g];
      for (int k = 0; k < 6; k++) {
        const double bk = bval(s_dsh[el][g][b], k, dd);
        if (bk == 0.0) continue;
        for (int h = 0; h < 6; h++) {
          const double bh = bval(s_dsh[el][g][a], h, d);
          const double c = s_c[el][g][h * 6 + k];
          if (bh == 0.0 || c == 0.0) continue;
          acc = acc + bh * c * bk * w;
        }
      }
    }
    ke_out[(long long)e * 576 + ent] = acc;
  }
}

// Affine displacement u = E X of Voigt strain s (micro/src/homogenize.c:335-342).
__device__ __forceinline__ double affine(const double *s, const double (&X)[3], int d)
{
#pragma clang fp contract(off)
  if (d == 0) return s[0] * X[0] + s[3] / 2 * X[1] + s[5] / 2 * X[2];
  if (d == 1) return s[3] / 2 * X[0] + s[1] * X[1] + s[4] / 2 * X[2];
  return s[5] / 2 * X[0] + s[4] / 2 * X[1] + s[2] * X[2];
}

// Element displacement: x at the element's unknowns, plus E X (true
// coordinates) for the periodic fluctuation.
__device__ void elem_disp(const FeDev &F, int e, const double *x, const double *strain, double (&ue)[24])
{
#pragma clang fp contract(off)
  for (int a = 0; a < 8; a++) {
    int ix, iy, iz;
    elem_node(F, e, a, ix, iy, iz);
    const long long u = unknown_of(F, ix, iy, iz);
    double X[3];
    if (F.periodic) node_coord(F, ix, iy, iz, X);
    for (int d = 0; d < 3; d++) ue[3 * a + d] = F.periodic ? x[3 * u + d] + affine(strain, X, d) : x[3 * u + d];
  }
}

// Gauss-point strain (all 24 terms, src order) and stress (C strain or given).
__device__ void gp_strain_stress(const FeDev &F, int e, int g, const double (&dsh)[8][3], const double (&ue)[24],
                                 const double *stress_gp, double (&eps)[6], double (&sig)[6])
{
#pragma clang fp contract(off)
  for (int v = 0; v < 6; v++) {
    double s = 0.0;
    for (int a = 0; a < 8; a++)
      for (int d = 0; d < 3; d++) s += bval(dsh[a], v, d) * ue[3 * a + d];
    eps[v] = (F.flush && fabs(s) < 1.0e-6) ? 0.0 : s;
  }
  if (stress_gp) {
    for (int v = 0; v < 6; v++) sig[v] = stress_gp[((long long)e * 8 + g) * 6 + v];
    return;
  }
  const double *c = F.c_gp ? F.c_gp + ((long long)e * 8 + g) * 36 : F.cmat + (size_t)F.elem_mat[e] * 36;
  for (int i = 0; i < 6; i++) {
    double s = 0.0;
    for (int j = 0; j < 6; j++) s += c[i * 6 + j] * eps[j];
    sig[i] = s;
  }
}

// Element residual res_e[i] = sum_gp sum_j B[j][i] sigma_j w
// (micro/src/assembly.c:45-56, macro/src/assembly.c:29-41); one lane per
// element.  With out_eps, the gauss-point strains are stored instead.
__global__ __launch_bounds__(256) void k_fe_res(FeDev F, const double *__restrict__ x, const double *strain,
                                               const double *__restrict__ stress_gp, double *__restrict__ res_out,
                                               double *__restrict__ eps_out, double *__restrict__ avg_out)
{
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= F.nelm) return;
  double ue[24], re[24], se[6], ss[6], wsum = 0.0;
  elem_disp(F, e, x, strain, ue);
  for (int i = 0; i < 24; i++) re[i] = 0.0;
  for (int v = 0; v < 6; v++) {
    se[v] = 0.0;
    ss[v] = 0.0;
  }
  for (int g = 0; g < 8; g++) {
    double dsh[8][3], w, eps[6], sig[6];
    gp_dsh(F, e, g, dsh, w);
    gp_strain_stress(F, e, g, dsh, ue, stress_gp, eps, sig);
    if (eps_out)
      for (int v = 0; v < 6; v++) eps_out[((long long)e * 8 + g) * 6 + v] = eps[v];
    if (res_out)
      for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++)
          for (int j = 0; j < 6; j++) re[3 * a + d] = re[3 * a + d] + bval(dsh[a], j, d) * sig[j] * w;
    for (int v = 0; v < 6; v++) {  // micro/src/homogenize.c:344-368
      ss[v] += sig[v] * w;
      se[v] += eps[v] * w;
    }
    wsum += w;
  }
  if (res_out)
    for (int i = 0; i < 24; i++) res_out[(long long)e * 24 + i] = re[i];
  if (avg_out) {  // element sums (for the volume average) and the element's weight
    double *o = avg_out + (long long)e * 13;
    for (int v = 0; v < 6; v++) {
      o[v] = se[v];
      o[6 + v] = ss[v];
    }
    o[12] = wsum;
  }
}

// The (element, local node) pairs of unknown node u in ascending element id
// (and local index), the order in which the reference's element loop
// reaches the node.
__device__ int node_elems(const FeDev &F, int u, int (&pe)[8], int (&pa)[8])
{
  int ux, uy, uz;
  if (F.periodic) {
    ux = u % F.nex;
    uy = (u / F.nex) % F.ney;
    uz = u / (F.nex * F.ney);
  } else {
    ux = u % F.nx;
    uy = (u / F.nx) % F.ny;
    uz = u / (F.nx * F.ny);
  }
  int n = 0;
  for (int a = 0; a < 8; a++) {
    int lx, ly, lz;
    local_coords(a, &lx, &ly, &lz);
    int ex = ux - lx, ey = uy - ly, ez = uz - lz;
    if (F.periodic) {
      ex = (ex + F.nex) % F.nex;
      ey = (ey + F.ney) % F.ney;
      ez = (ez + F.nez) % F.nez;
    } else if (ex < 0 || ey < 0 || ez < 0 || ex >= F.nex || ey >= F.ney || ez >= F.nez) {
      continue;
    }
    pe[n] = ex + ey * F.nex + ez * F.nex * F.ney;
    pa[n] = a;
    n++;
  }
  for (int i = 1; i < n; i++)  // insertion sort by (element, local node)
    for (int j = i; j > 0 && (pe[j] < pe[j - 1] || (pe[j] == pe[j - 1] && pa[j] < pa[j - 1])); j--) {
      const int te = pe[j], ta = pa[j];
      pe[j] = pe[j - 1];
      pa[j] = pa[j - 1];
      pe[j - 1] = te;
      pa[j - 1] = ta;
    }
  return n;
}

__device__ __forceinline__ int elem_unknown(const FeDev &F, int e, int b)
{
  int ix, iy, iz;
  elem_node(F, e, b, ix, iy, iz);
  return unknown_of(F, ix, iy, iz);
}

// Residual of unknown node u: res[3u+d] = sum over its (element, a) pairs in
// order of res_e[3a+d] (micro/src/assembly.c:57); Dirichlet dofs 0.
// Partial ||res||^2 per workgroup.
__global__ __launch_bounds__(256) void k_fe_node_res(FeDev F, const double *__restrict__ res_e,
                                                    double *__restrict__ res, double *__restrict__ part)
{
#pragma clang fp contract(off)
  __shared__ double lds[16];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (u < F.nunk) {
    int pe[8], pa[8];
    const int n = node_elems(F, u, pe, pa);
    for (int d = 0; d < 3; d++) {
      double r = 0.0;
      for (int c = 0; c < n; c++) r = r + res_e[(long long)pe[c] * 24 + 3 * pa[c] + d];
      if (F.dir[3 * (long long)u + d]) r = 0.0;
      res[3 * (long long)u + d] = r;
      v[0] += r * r;
    }
  }
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// One workgroup sums n partial values of width w (value i of item k at
// p[k * w + i]) in item order; out[i] receives the totals.
__global__ __launch_bounds__(256) void k_fe_sum(const double *__restrict__ p, int n, int w, double *__restrict__ out)
{
  __shared__ double lds[16 * 13];
  double s[13];
  for (int i = 0; i < 13; i++) s[i] = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x)
    for (int i = 0; i < w; i++) s[i] += p[(long long)k * w + i];
  block_sum<13>(s, lds);
  if (threadIdx.x == 0)
    for (int i = 0; i < w; i++) out[i] = s[i];
}

// ELL rows of unknown node u in the reference's layout: the neighbours in
// first-touch order (elements ascending, then their local nodes), each
// slot the sum of the contributing element entries in ascending element
// order (src/ell.c:46-84 over micro/src/assembly.c:93-110).  Dirichlet rows
// become identity rows with their columns kept (src/ell.c:86-101); with
// zero_cols, Dirichlet columns are zeroed in the other rows.  Output in the
// slab layout of the device ELL (S rows per slab, column-major slots), and
// dinv = 1 / diagonal.
__global__ __launch_bounds__(256) void k_fe_node_ell(FeDev F, const double *__restrict__ ke, int S,
                                                    int *__restrict__ cols, double *__restrict__ vals,
                                                    double *__restrict__ dinv)
{
#pragma clang fp contract(off)
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= F.nunk) return;
  int pe[8], pa[8];
  const int np = node_elems(F, u, pe, pa);
  int nb[27], nnb = 0;
  for (int c = 0; c < np; c++) {
    if (c > 0 && pe[c] == pe[c - 1]) continue;  // same element again: no new columns
    for (int b = 0; b < 8; b++) {
      const int m = elem_unknown(F, pe[c], b);
      bool seen = false;
      for (int t = 0; t < nnb; t++) seen = seen || nb[t] == m;
      if (!seen && nnb < 27) nb[nnb++] = m;
    }
  }
  for (int d = 0; d < 3; d++) {
    const long long row = 3LL * u + d;
    const long long base = (row / S) * (long long)S * kNnz + row % S;
    const bool drow = F.dir[row];
    double diag = 0.0;
    for (int t = 0; t < 27; t++)
      for (int dd = 0; dd < 3; dd++) {
        const int j = 3 * t + dd;
        int col = -1;
        double v = 0.0;
        if (t < nnb) {
          col = 3 * nb[t] + dd;
          bool first = true;
          for (int c = 0; c < np; c++)
            for (int b = 0; b < 8; b++) {
              if (elem_unknown(F, pe[c], b) != nb[t]) continue;
              const double kv = ke[(long long)pe[c] * 576 + (3 * pa[c] + d) * 24 + 3 * b + dd];
              v = first ? kv : v + kv;
              first = false;
            }
          if (drow) v = (col == row) ? 1.0 : 0.0;
          else if (F.zero_cols && F.dir[col]) v = 0.0;
          if (col == row) diag = v;
        }
        cols[base + (long long)j * S] = col;
        vals[base + (long long)j * S] = v;
      }
    dinv[row] = 1 / diag;
  }
}

// x[dof] = E X on the Dirichlet dofs (set_disp_0).
__global__ void k_fe_disp0(FeDev F, const double *strain, double *__restrict__ x)
{
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= F.nunk || F.periodic) return;
  const int ix = u % F.nx, iy = (u / F.nx) % F.ny, iz = u / (F.nx * F.ny);
  double X[3];
  node_coord(F, ix, iy, iz, X);
  for (int d = 0; d < 3; d++)
    if (F.dir[3LL * u + d]) x[3LL * u + d] = affine(strain, X, d);
}

// b = -r
__global__ void k_fe_neg(const double *__restrict__ r, double *__restrict__ b, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = -r[i];
}

// x += dx
__global__ void k_fe_add(double *__restrict__ x, const double *__restrict__ dx, int n)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + dx[i];
}

__global__ void k_fe_zero(double *__restrict__ x, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

}  // namespace

struct Fe {
  FeDev F;
  int device = 0;
  hipStream_t s = nullptr;
  ell_device *A = nullptr;
  int *ell_cols = nullptr;
  double *ell_vals = nullptr, *ell_dinv = nullptr;
  int S = 256;
  double *coords = nullptr, *cmat = nullptr, *c_gp = nullptr;
  uint8_t *elem_mat = nullptr, *dir = nullptr;
  double *x = nullptr, *res = nullptr, *b = nullptr, *dx = nullptr;
  double *ke = nullptr, *res_e = nullptr, *part = nullptr, *tot = nullptr, *strain = nullptr;
  double *stress_gp = nullptr, *gp_buf = nullptr, *avg = nullptr;
  std::vector<uint8_t> phase;
};

static int grid_of(long long n, int bs) { return (int)((n + bs - 1) / bs); }

int fe_create(const FeMeshDesc &d, Fe **out)
{
  if (!out || d.nex < 1 || d.ney < 1 || d.nez < 1 || !(d.lx > 0) || !(d.ly > 0) || !(d.lz > 0)) return MM_ERR_ARG;
  if (d.nmat < 1 || !d.cmat) return MM_ERR_ARG;
  if (d.periodic && (d.nex < 3 || d.ney < 3 || d.nez < 3 || d.coords)) return MM_ERR_ARG;
  const long long nn = (long long)(d.nex + 1) * (d.ney + 1) * (d.nez + 1);
  const long long nunk = d.periodic ? (long long)d.nex * d.ney * d.nez : nn;
  const long long nelm = (long long)d.nex * d.ney * d.nez;
  if (3 * nunk > 2000000000LL || nelm * 576 > (1LL << 40)) return MM_ERR_ARG;
  MM_HIP(hipSetDevice(d.device));
  Fe *f = new Fe();
  f->device = d.device;
  FeDev &F = f->F;
  memset(&F, 0, sizeof(F));
  F.nex = d.nex;
  F.ney = d.ney;
  F.nez = d.nez;
  F.nx = d.nex + 1;
  F.ny = d.ney + 1;
  F.nz = d.nez + 1;
  F.nelm = (int)nelm;
  F.nunk = (int)nunk;
  F.nrow = (int)(3 * nunk);
  F.periodic = d.periodic;
  F.iso = d.coords ? 1 : 0;
  F.flush = d.flush_strain;
  F.hx = d.lx / d.nex;
  F.hy = d.ly / d.ney;
  F.hz = d.lz / d.nez;
  F.vol = d.lx * d.ly * d.lz;
  f->phase.assign((size_t)nelm, 0);
  for (long long e = 0; e < nelm; e++) {
    const int m = d.elem_mat ? d.elem_mat[e] : 0;
    if (m < 0 || m >= d.nmat) {
      delete f;
      return MM_ERR_ARG;
    }
    f->phase[e] = (uint8_t)m;
  }
  int rc = 0;
#define FALLOC(ptr, bytes)                                          \
  do {                                                              \
    if (hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 8) != hipSuccess) { \
      rc = MM_ERR_ALLOC;                                            \
      goto fail;                                                    \
    }                                                               \
  } while (0)
  if (hipStreamCreateWithFlags(&f->s, hipStreamNonBlocking) != hipSuccess) {
    delete f;
    return MM_ERR_HIP;
  }
  FALLOC(f->elem_mat, (size_t)nelm);
  FALLOC(f->cmat, (size_t)d.nmat * 36 * 8);
  FALLOC(f->dir, (size_t)F.nrow);
  FALLOC(f->x, (size_t)F.nrow * 8);
  FALLOC(f->res, (size_t)F.nrow * 8);
  FALLOC(f->b, (size_t)F.nrow * 8);
  FALLOC(f->dx, (size_t)F.nrow * 8);
  FALLOC(f->ke, (size_t)nelm * 576 * 8);
  FALLOC(f->res_e, (size_t)nelm * 24 * 8);
  FALLOC(f->avg, (size_t)nelm * 13 * 8);
  FALLOC(f->part, (size_t)(grid_of(nunk, 256) + 1) * 8);
  FALLOC(f->tot, 16 * 8);
  FALLOC(f->strain, 8 * 8);
  if (d.coords) {
    FALLOC(f->coords, (size_t)nn * 3 * 8);
    MM_HIP(hipMemcpyAsync(f->coords, d.coords, (size_t)nn * 3 * 8, hipMemcpyHostToDevice, f->s));
  }
#undef FALLOC
  if (hipMemcpyAsync(f->elem_mat, f->phase.data(), (size_t)nelm, hipMemcpyHostToDevice, f->s) != hipSuccess ||
      hipMemcpyAsync(f->cmat, d.cmat, (size_t)d.nmat * 36 * 8, hipMemcpyHostToDevice, f->s) != hipSuccess ||
      hipMemsetAsync(f->dir, 0, (size_t)F.nrow, f->s) != hipSuccess ||
      hipMemsetAsync(f->x, 0, (size_t)F.nrow * 8, f->s) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  if ((rc = ell_device_alloc(F.nrow, kNnz, 256, &f->A))) goto fail;
  {
    hipStream_t as;
    ell_device_arrays(f->A, &f->ell_cols, &f->ell_vals, &f->ell_dinv, &f->S, &as);
  }
  F.coords = f->coords;
  F.elem_mat = f->elem_mat;
  F.cmat = f->cmat;
  F.c_gp = nullptr;
  F.dir = f->dir;
  if (F.periodic) {  // the fluctuation of node (0,0,0) is pinned (all corners get u = E X)
    const uint8_t one[3] = {1, 1, 1};
    if (hipMemcpyAsync(f->dir, one, 3, hipMemcpyHostToDevice, f->s) != hipSuccess) {
      rc = MM_ERR_HIP;
      goto fail;
    }
    F.zero_cols = 1;
  }
  if (hipStreamSynchronize(f->s) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  *out = f;
  return 0;
fail:
  fe_destroy(f);
  return rc ? rc : MM_ERR_HIP;
}

void fe_destroy(Fe *f)
{
  if (!f) return;
  if (f->s) (void)hipStreamSynchronize(f->s);
  void *p[] = {f->coords, f->cmat, f->c_gp, f->elem_mat, f->dir, f->x, f->res, f->b, f->dx, f->ke,
               f->res_e, f->part, f->tot, f->strain, f->stress_gp, f->gp_buf, f->avg};
  for (void *q : p)
    if (q) (void)hipFree(q);
  if (f->A) ell_device_destroy(f->A);
  if (f->s) (void)hipStreamDestroy(f->s);
  delete f;
}

int fe_sizes(const Fe *f, int *nrow, int *nelm, int *nnode)
{
  if (!f) return MM_ERR_ARG;
  if (nrow) *nrow = f->F.nrow;
  if (nelm) *nelm = f->F.nelm;
  if (nnode) *nnode = f->F.nunk;
  return 0;
}

int fe_set_dirichlet(Fe *f, int ndof, const int *dofs, int zero_cols)
{
  if (!f || ndof < 0 || (ndof && !dofs)) return MM_ERR_ARG;
  std::vector<uint8_t> h((size_t)f->F.nrow, 0);
  for (int i = 0; i < ndof; i++) {
    if (dofs[i] < 0 || dofs[i] >= f->F.nrow) return MM_ERR_ARG;
    h[dofs[i]] = 1;
  }
  if (f->F.periodic) h[0] = h[1] = h[2] = 1;
  MM_HIP(hipMemcpyAsync(f->dir, h.data(), h.size(), hipMemcpyHostToDevice, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  f->F.zero_cols = (zero_cols || f->F.periodic) ? 1 : 0;
  return 0;
}

int fe_set_dirichlet_boundary(Fe *f)
{
  if (!f || f->F.periodic) return MM_ERR_ARG;
  std::vector<int> dofs;
  const FeDev &F = f->F;
  for (int u = 0; u < F.nunk; u++) {
    const int ix = u % F.nx, iy = (u / F.nx) % F.ny, iz = u / (F.nx * F.ny);
    if (ix == 0 || iy == 0 || iz == 0 || ix == F.nx - 1 || iy == F.ny - 1 || iz == F.nz - 1)
      for (int d = 0; d < 3; d++) dofs.push_back(3 * u + d);
  }
  return fe_set_dirichlet(f, (int)dofs.size(), dofs.data(), 0);
}

int fe_set_x(Fe *f, const double *x)
{
  if (!f || !x) return MM_ERR_ARG;
  MM_HIP(hipMemcpyAsync(f->x, x, (size_t)f->F.nrow * 8, hipMemcpyHostToDevice, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  return 0;
}

int fe_get_x(Fe *f, double *x)
{
  if (!f || !x) return MM_ERR_ARG;
  MM_HIP(hipMemcpyAsync(x, f->x, (size_t)f->F.nrow * 8, hipMemcpyDeviceToHost, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  return 0;
}

static int upload_strain(Fe *f, const double *strain)
{
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (strain)
    for (int i = 0; i < 6; i++) s[i] = strain[i];
  MM_HIP(hipMemcpyAsync(f->strain, s, sizeof(s), hipMemcpyHostToDevice, f->s));
  return 0;
}

int fe_set_disp_0(Fe *f, const double *strain)
{
  if (!f || !strain) return MM_ERR_ARG;
  int rc = upload_strain(f, strain);
  if (rc) return rc;
  k_fe_disp0<<<grid_of(f->F.nunk, 256), 256, 0, f->s>>>(f->F, f->strain, f->x);
  MM_HIP(hipGetLastError());
  MM_HIP(hipStreamSynchronize(f->s));
  return 0;
}

int fe_set_c_gp(Fe *f, const double *c_gp)
{
  if (!f) return MM_ERR_ARG;
  if (!c_gp) {
    f->F.c_gp = nullptr;
    return 0;
  }
  if (!f->c_gp && hipMalloc((void **)&f->c_gp, (size_t)f->F.nelm * 8 * 36 * 8) != hipSuccess) return MM_ERR_ALLOC;
  MM_HIP(hipMemcpyAsync(f->c_gp, c_gp, (size_t)f->F.nelm * 8 * 36 * 8, hipMemcpyHostToDevice, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  f->F.c_gp = f->c_gp;
  return 0;
}

static int upload_stress(Fe *f, const double *stress_gp, const double **dev)
{
  *dev = nullptr;
  if (!stress_gp) return 0;
  if (!f->stress_gp && hipMalloc((void **)&f->stress_gp, (size_t)f->F.nelm * 8 * 6 * 8) != hipSuccess)
    return MM_ERR_ALLOC;
  MM_HIP(hipMemcpyAsync(f->stress_gp, stress_gp, (size_t)f->F.nelm * 8 * 6 * 8, hipMemcpyHostToDevice, f->s));
  *dev = f->stress_gp;
  return 0;
}

int fe_residual(Fe *f, const double *strain, const double *stress_gp, double *norm)
{
  if (!f) return MM_ERR_ARG;
  int rc = upload_strain(f, strain);
  if (rc) return rc;
  const double *sg = nullptr;
  if ((rc = upload_stress(f, stress_gp, &sg))) return rc;
  const FeDev &F = f->F;
  k_fe_res<<<grid_of(F.nelm, 256), 256, 0, f->s>>>(F, f->x, f->strain, sg, f->res_e, nullptr, nullptr);
  const int g = grid_of(F.nunk, 256);
  k_fe_node_res<<<g, 256, 0, f->s>>>(F, f->res_e, f->res, f->part);
  k_fe_sum<<<1, 256, 0, f->s>>>(f->part, g, 1, f->tot);
  MM_HIP(hipGetLastError());
  double t = 0.0;
  MM_HIP(hipMemcpyAsync(&t, f->tot, 8, hipMemcpyDeviceToHost, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  if (norm) *norm = std::sqrt(t);
  return 0;
}

int fe_get_residual(Fe *f, double *r)
{
  if (!f || !r) return MM_ERR_ARG;
  MM_HIP(hipMemcpyAsync(r, f->res, (size_t)f->F.nrow * 8, hipMemcpyDeviceToHost, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  return 0;
}

int fe_gp_strain(Fe *f, const double *strain, double *out)
{
  if (!f || !out) return MM_ERR_ARG;
  int rc = upload_strain(f, strain);
  if (rc) return rc;
  if (!f->gp_buf && hipMalloc((void **)&f->gp_buf, (size_t)f->F.nelm * 8 * 6 * 8) != hipSuccess)
    return MM_ERR_ALLOC;
  // the stress of the gauss points is computed and discarded: with a
  // stress input absent the tangent supplies it
  k_fe_res<<<grid_of(f->F.nelm, 256), 256, 0, f->s>>>(f->F, f->x, f->strain, nullptr, nullptr, f->gp_buf, nullptr);
  MM_HIP(hipGetLastError());
  MM_HIP(hipMemcpyAsync(out, f->gp_buf, (size_t)f->F.nelm * 8 * 6 * 8, hipMemcpyDeviceToHost, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  return 0;
}

int fe_assemble(Fe *f)
{
  if (!f) return MM_ERR_ARG;
  const FeDev &F = f->F;
  k_fe_ke<<<grid_of(F.nelm, kEB), 64, 0, f->s>>>(F, f->ke);
  k_fe_node_ell<<<grid_of(F.nunk, 256), 256, 0, f->s>>>(F, f->ke, f->S, f->ell_cols, f->ell_vals, f->ell_dinv);
  MM_HIP(hipGetLastError());
  MM_HIP(hipStreamSynchronize(f->s));
  return 0;
}

int fe_get_ell(Fe *f, int *cols, double *vals)
{
  if (!f || !cols || !vals) return MM_ERR_ARG;
  return ell_device_download(f->A, cols, vals);
}

int fe_solve_update(Fe *f, double rtol, int max_its, int *its)
{
  if (!f) return MM_ERR_ARG;
  const int n = f->F.nrow;
  k_fe_neg<<<grid_of(n, 256), 256, 0, f->s>>>(f->res, f->b, n);
  k_fe_zero<<<grid_of(n, 256), 256, 0, f->s>>>(f->dx, n);
  MM_HIP(hipGetLastError());
  MM_HIP(hipStreamSynchronize(f->s));
  ell_solver sv = {max_its, 0, rtol, 0.0};
  const int rc = ell_device_pcg_device(f->A, &sv, f->b, f->dx);
  if (its) *its = sv.its;
  if (rc && rc != MM_NOT_CONVERGED) return rc;
  k_fe_add<<<grid_of(n, 256), 256, 0, f->s>>>(f->x, f->dx, n);
  MM_HIP(hipGetLastError());
  MM_HIP(hipStreamSynchronize(f->s));
  return rc;
}

int fe_solve(Fe *f, const double *b, double *du, double rtol, int max_its, int *its)
{
  if (!f || !b || !du) return MM_ERR_ARG;
  const int n = f->F.nrow;
  MM_HIP(hipMemcpyAsync(f->b, b, (size_t)n * 8, hipMemcpyHostToDevice, f->s));
  k_fe_zero<<<grid_of(n, 256), 256, 0, f->s>>>(f->dx, n);
  MM_HIP(hipGetLastError());
  MM_HIP(hipStreamSynchronize(f->s));
  ell_solver sv = {max_its, 0, rtol, 0.0};
  const int rc = ell_device_pcg_device(f->A, &sv, f->b, f->dx);
  if (its) *its = sv.its;
  if (rc && rc != MM_NOT_CONVERGED) return rc;
  MM_HIP(hipMemcpyAsync(du, f->dx, (size_t)n * 8, hipMemcpyDeviceToHost, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  return rc;
}

int fe_averages(Fe *f, const double *strain, double *strain_ave, double *stress_ave, double *elem_strain,
                double *elem_stress)
{
  if (!f) return MM_ERR_ARG;
  int rc = upload_strain(f, strain);
  if (rc) return rc;
  const FeDev &F = f->F;
  k_fe_res<<<grid_of(F.nelm, 256), 256, 0, f->s>>>(F, f->x, f->strain, nullptr, nullptr, nullptr, f->avg);
  k_fe_sum<<<1, 256, 0, f->s>>>(f->avg, F.nelm, 13, f->tot);
  MM_HIP(hipGetLastError());
  double t[16];
  MM_HIP(hipMemcpyAsync(t, f->tot, 13 * 8, hipMemcpyDeviceToHost, f->s));
  MM_HIP(hipStreamSynchronize(f->s));
  for (int v = 0; v < 6; v++) {
    if (strain_ave) strain_ave[v] = t[v] / F.vol;
    if (stress_ave) stress_ave[v] = t[6 + v] / F.vol;
  }
  if (elem_strain || elem_stress) {
    std::vector<double> h((size_t)F.nelm * 13);
    MM_HIP(hipMemcpyAsync(h.data(), f->avg, h.size() * 8, hipMemcpyDeviceToHost, f->s));
    MM_HIP(hipStreamSynchronize(f->s));
    for (int e = 0; e < F.nelm; e++)
      for (int v = 0; v < 6; v++) {
        const double *o = &h[(size_t)e * 13];
        if (elem_strain) elem_strain[(size_t)e * 6 + v] = o[v] / o[12];
        if (elem_stress) elem_stress[(size_t)e * 6 + v] = o[6 + v] / o[12];
      }
  }
  return 0;
}

int fe_get_phase(Fe *f, int *elem_type)
{
  if (!f || !elem_type) return MM_ERR_ARG;
  for (int e = 0; e < f->F.nelm; e++) elem_type[e] = f->phase[e];
  return 0;
}

}  // namespace mm
