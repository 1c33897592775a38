// micro.cpp -- the reference's process-global micro homogenization entry
// points (micro/inc/micro.h:150-157, micro/src/homogenize.c) served by a
// one-RVE engine.  The RVE stays resident in HBM between calls (assembled
// once); every solve starts from zero interior displacement.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "fe.h"
#include "micro_api.h"

namespace {

constexpr uint64_t kKey = 1;  // resident RVE key inside the private engine

struct MicroGlobal {
  int init = 0;
  micro_config cfg;
  mm_engine *eng = nullptr;
  double c_tangent_linear[36];  // micro/inc/micro.h:90
  double c_tangent[36];         // micro/inc/micro.h:91
  bool linear_materials = true;      // flags.linear_materials (micro/inc/micro.h:100)
  bool c_linear_calculated = false;  // micro/inc/micro.h:101
  // element path: per-iteration re-assembly (linear_materials == 0) and the
  // periodic boundary condition run on the element assembly of fe.hip
  mm::Fe *fe = nullptr;
  std::vector<double> last_strain;  // macro strain of the last homog_fe2
  micro_stats st;
  uint64_t next_id = 1;
} g;

// Phase map of the RVE (micro/src/micro_struct.c:114-127), as the GPU's
// k_phase computes it.
void phase_map(const mm_rve_spec &r, std::vector<int> &et)
{
#pragma clang fp contract(off)
  const int nelm = r.nex * r.ney * r.nez;
  const double hx = r.lx / r.nex, hy = r.ly / r.ney;
  et.assign(nelm, 0);
  for (int e = 0; e < nelm; e++) {
    const int ex = e % r.nex, ey = (e / r.nex) % r.ney;
    const double cx = (ex + 0.5) * hx, cy = (ey + 0.5) * hy;
    for (int f = 0; f < r.nfib; f++) {
      const double dx = cx - r.fib_x[f], dy = cy - r.fib_y[f];
      double l = 0.0;
      l = l + dx * dx;
      l = l + dy * dy;
      l = std::sqrt(l);
      if (l <= r.fib_r[f]) {
        et[e] = 1;
        break;
      }
    }
  }
}

// Isotropic tangent in the form of src/material.c:16-22.
void iso_c(double E, double nu, double *c)
{
#pragma clang fp contract(off)
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      double base = 0.0;
      if (i < 3 && j < 3) base = (i == j) ? 1.0 - nu : nu;
      else if (i == j) base = (1 - 2 * nu) / 2;
      c[i * 6 + j] = base * (E / ((1 + nu) * (1 - 2 * nu)));
    }
}

int fe_init(const micro_config *cfg)
{
  std::vector<int> et;
  phase_map(cfg->rve, et);
  double cmat[2 * 36];
  for (int m = 0; m < 2; m++) iso_c(cfg->rve.young[m], cfg->rve.poisson[m], cmat + 36 * m);
  mm::FeMeshDesc d;
  memset(&d, 0, sizeof(d));
  d.nex = cfg->rve.nex;
  d.ney = cfg->rve.ney;
  d.nez = cfg->rve.nez;
  d.lx = cfg->rve.lx;
  d.ly = cfg->rve.ly;
  d.lz = cfg->rve.lz;
  d.periodic = (cfg->fe2_bc == MICRO_BC_PER_LM) ? 1 : 0;
  d.elem_mat = et.data();
  d.nmat = 2;
  d.cmat = cmat;
  d.device = cfg->device;
  int rc = mm::fe_create(d, &g.fe);
  if (rc) return rc;
  if (!d.periodic) rc = mm::fe_set_dirichlet_boundary(g.fe);
  return rc;
}

// homog_fe2 on the element path (micro/src/homogenize.c:175-225): from zero
// interior displacement (set_disp_0 gives the boundary values), while
// nl_its < nl_max_its and |r| > nl_min_norm: residual, then the Jacobian
// assembled again from the gauss-point tangents, a Jacobi-PCG solve of
// K dx = -r and x += dx; then the volume averages.
int fe_homog(const double *strain, double *strain_ave, double *stress_ave)
{
  const auto t0 = std::chrono::steady_clock::now();
  int rc = 0;
  int nrow = 0;
  mm::fe_sizes(g.fe, &nrow, nullptr, nullptr);
  std::vector<double> zero((size_t)nrow, 0.0);
  if ((rc = mm::fe_set_x(g.fe, zero.data()))) return rc;
  if (g.cfg.fe2_bc != MICRO_BC_PER_LM && (rc = mm::fe_set_disp_0(g.fe, strain))) return rc;
  int nl = 0, its_total = 0;
  double norm = g.cfg.nl_min_norm * 10;
  while (nl < g.cfg.nl_max_its && norm > g.cfg.nl_min_norm) {
    if ((rc = mm::fe_residual(g.fe, strain, nullptr, &norm))) return rc;
    if (norm < g.cfg.nl_min_norm) break;
    if ((rc = mm::fe_assemble(g.fe))) return rc;
    g.st.assemblies++;
    int its = 0;
    rc = mm::fe_solve_update(g.fe, g.cfg.rtol, 100000, &its);
    its_total += its;
    if (rc) return rc;
    nl++;
  }
  if ((rc = mm::fe_averages(g.fe, strain, strain_ave, stress_ave, nullptr, nullptr))) return rc;
  g.last_strain.assign(strain, strain + MM_NVOI);
  g.st.pcg_its = its_total;
  g.st.sweeps = nl;
  g.st.residual = norm;
  g.st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int run_request(const double (*loads)[MM_NVOI], int nload, int tangent_mode, mm_result *res)
{
  mm_request rq;
  memset(&rq, 0, sizeof(rq));
  rq.id = g.next_id++;
  rq.rve = g.cfg.rve;
  rq.nload = nload;
  for (int k = 0; k < nload; k++)
    for (int i = 0; i < MM_NVOI; i++) rq.strain[k][i] = loads[k][i];
  rq.tangent_mode = tangent_mode;
  rq.warm_key = kKey;
  rq.warm_start = 0;
  g.last_strain.assign(loads[0], loads[0] + MM_NVOI);
  const auto t0 = std::chrono::steady_clock::now();
  int rc = mm_engine_submit(g.eng, &rq);
  if (rc) return rc;
  rc = mm_engine_drain(g.eng);
  if (rc) return rc;
  int n = 0;
  rc = mm_engine_poll(g.eng, res, 1, &n);
  if (rc) return rc;
  if (n != 1 || res->id != rq.id) return MM_ERR_STATE;
  g.st.pcg_its = res->pcg_its;
  g.st.sweeps = res->sweeps;
  g.st.residual = res->residual_rel;
  g.st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace

extern "C" int micro_config_default(micro_config *c)
{
  if (!c) return MM_ERR_ARG;
  memset(c, 0, sizeof(*c));
  c->rve.nex = c->rve.ney = c->rve.nez = 10;
  c->rve.lx = c->rve.ly = c->rve.lz = 1.0;
  c->rve.nfib = 1;
  c->rve.fib_x[0] = 0.5;
  c->rve.fib_y[0] = 0.5;
  c->rve.fib_r[0] = 0.25;
  c->rve.young[0] = 1.0e7;
  c->rve.poisson[0] = 0.25;
  c->rve.young[1] = 1.0e8;
  c->rve.poisson[1] = 0.3;
  c->solver = MICRO_SOLVER_PCG;
  c->rtol = 1.0e-10;
  c->nl_max_its = 2;     // micro/src/init.c:7
  c->nl_min_norm = 1.0e-4;  // micro/src/init.c:8
  c->rho = 1.0e7;           // micro/src/init.c:9
  c->device = 0;
  c->linear_materials = 1;
  c->fe2_bc = MICRO_BC_USTRAIN;  // micro/src/init.c default
  return 0;
}

extern "C" int micro_finalize(void)
{
  if (g.eng) mm_engine_destroy(g.eng);
  if (g.fe) mm::fe_destroy(g.fe);
  g = MicroGlobal();
  return 0;
}

extern "C" int homog_calculate_c_tangent(double *strain_mac, double *c_tangent);

// micro/src/main.c:57-109 and init_variables_2 (micro/src/init.c:37-62)
extern "C" int micro_init(const micro_config *cfg)
{
  if (!cfg) return MM_ERR_ARG;
  if (g.init) micro_finalize();
  if (cfg->solver != MICRO_SOLVER_PCG && cfg->solver != MICRO_SOLVER_PCG_MIXED) return MM_ERR_ARG;
  if (!(cfg->rtol > 0) || cfg->nl_max_its < 1) return MM_ERR_ARG;
  if (cfg->linear_materials != 0 && cfg->linear_materials != 1) return MM_ERR_ARG;
  if (cfg->fe2_bc == MICRO_BC_USTRESS) {
    fprintf(stderr, "madmax_amd: BC_USTRESS has no assembly in the reference (micro/src/alloc.c:18-19 only "
                    "sizes it)\n");
    return MM_ERR_ARG;
  }
  if (cfg->fe2_bc != MICRO_BC_USTRAIN && cfg->fe2_bc != MICRO_BC_PER_LM) return MM_ERR_ARG;
  g.cfg = *cfg;
  g.linear_materials = cfg->linear_materials != 0;
  if (!g.linear_materials || cfg->fe2_bc == MICRO_BC_PER_LM) {
    // element path: FP64 Newton with per-iteration assembly
    int rc = fe_init(cfg);
    if (rc) {
      micro_finalize();
      return rc;
    }
    g.init = 1;
    g.c_linear_calculated = false;
    if (g.linear_materials) {
      rc = homog_calculate_c_tangent_around_zero(g.c_tangent_linear);
      if (rc) return rc;
      g.c_linear_calculated = true;
    }
    return 0;
  }
  mm_engine_config ec;
  mm_engine_config_default(&ec);
  ec.device = cfg->device;
  ec.max_rves = 2;
  ec.slab_nodes = 128;
  ec.nrhs = 7;
  ec.use_graph = 1;
  ec.steps_per_launch = 32;
  if (cfg->solver == MICRO_SOLVER_PCG) {
    // Newton loop of homog_fe2 with a FP64 PCG solve per iteration, from the
    // reference's start (boundary values, interior zero)
    ec.precision = MM_PREC_FP64;
    ec.initial_guess = MM_GUESS_ZERO;
    ec.rtol = 0.0;
    ec.abs_tol = cfg->nl_min_norm;
    ec.inner_rtol = cfg->rtol;
    ec.max_inner_its = 100000;
    ec.max_sweeps = cfg->nl_max_its;
  } else {
    ec.precision = MM_PREC_MIXED;
    ec.rtol = cfg->rtol;
    ec.abs_tol = cfg->nl_min_norm;
    ec.inner_rtol = 1.0e-5;
    ec.max_inner_its = 100000;
    ec.max_sweeps = cfg->nl_max_its > 12 ? cfg->nl_max_its : 12;
  }
  const double rows = 3.0 * (cfg->rve.nex + 1.0) * (cfg->rve.ney + 1.0) * (cfg->rve.nez + 1.0);
  ec.hbm_budget = (size_t)((rows + 6.0 * ec.slab_nodes) * 1800.0) + (size_t)(8u << 20);
  int rc = mm_engine_create(&ec, &g.eng);
  if (rc) return rc;
  g.init = 1;
  g.c_linear_calculated = false;
  if (g.linear_materials) {
    rc = homog_calculate_c_tangent_around_zero(g.c_tangent_linear);
    if (rc) return rc;
    g.c_linear_calculated = true;
  }
  return 0;
}

// micro/src/homogenize.c:175-225
extern "C" int homog_fe2(double *strain_mac, double *strain_ave, double *stress_ave)
{
  if (!g.init) return MM_ERR_STATE;
  if (!strain_mac || !strain_ave || !stress_ave) return MM_ERR_ARG;
  if (g.fe) return fe_homog(strain_mac, strain_ave, stress_ave);
  double load[1][MM_NVOI];
  for (int i = 0; i < MM_NVOI; i++) load[0][i] = strain_mac[i];
  mm_result res;
  int rc = run_request(load, 1, MM_TANGENT_NONE, &res);
  if (rc) return rc;
  for (int i = 0; i < MM_NVOI; i++) {
    strain_ave[i] = res.strain[0][i];
    stress_ave[i] = res.stress[0][i];
  }
  return 0;
}

// micro/src/homogenize.c:23-33
extern "C" int homog_get_strain_stress_non_linear(double *strain_mac, double *strain_ave,
                                                  double *stress_ave)
{
  return homog_fe2(strain_mac, strain_ave, stress_ave);
}

// micro/src/homogenize.c:3-21
extern "C" int homog_get_strain_stress(double *strain_mac, double *strain_ave, double *stress_ave)
{
  if (!g.init) return MM_ERR_STATE;
  if (!strain_mac || !strain_ave || !stress_ave) return MM_ERR_ARG;
  if (g.linear_materials && g.c_linear_calculated) {
    for (int i = 0; i < MM_NVOI; i++) {
      strain_ave[i] = strain_mac[i];
      stress_ave[i] = 0.0;
      for (int j = 0; j < MM_NVOI; j++) stress_ave[i] += g.c_tangent_linear[i * MM_NVOI + j] * strain_mac[j];
    }
    return 0;
  }
  return homog_get_strain_stress_non_linear(strain_mac, strain_ave, stress_ave);
}

// micro/src/homogenize.c:59-94: sigma(eps) and sigma(eps + delta e_i) as one
// 7-load-case block.  Like the reference, once the linear tangent is known
// the stresses come from it (homog_get_strain_stress takes the shortcut).
extern "C" int homog_calculate_c_tangent(double *strain_mac, double *c_tangent)
{
  if (!g.init) return MM_ERR_STATE;
  if (!strain_mac || !c_tangent) return MM_ERR_ARG;
  double loads[7][MM_NVOI];
  for (int k = 0; k < 7; k++)
    for (int i = 0; i < MM_NVOI; i++) loads[k][i] = strain_mac[i];
  for (int i = 0; i < MM_NVOI; i++) loads[i + 1][i] = loads[i + 1][i] + HOMOGENIZE_DELTA_STRAIN;
  if (g.linear_materials && g.c_linear_calculated) {
    double s1[MM_NVOI], s2[MM_NVOI], aux[MM_NVOI];
    homog_get_strain_stress(loads[0], aux, s1);
    for (int i = 0; i < MM_NVOI; i++) {
      homog_get_strain_stress(loads[i + 1], aux, s2);
      for (int j = 0; j < MM_NVOI; j++)
        c_tangent[j * MM_NVOI + i] = (s2[j] - s1[j]) / (loads[i + 1][i] - loads[0][i]);
    }
    return 0;
  }
  if (g.fe) {  // seven homog_fe2 solves, as micro/src/homogenize.c:59-94
    double s1[MM_NVOI], s2[MM_NVOI], aux[MM_NVOI];
    int rc = homog_fe2(loads[0], aux, s1);
    if (rc) return rc;
    for (int i = 0; i < MM_NVOI; i++) {
      if ((rc = homog_fe2(loads[i + 1], aux, s2))) return rc;
      for (int j = 0; j < MM_NVOI; j++)
        c_tangent[j * MM_NVOI + i] = (s2[j] - s1[j]) / (loads[i + 1][i] - loads[0][i]);
    }
    return 0;
  }
  mm_result res;
  int rc = run_request(loads, 7, MM_TANGENT_FD, &res);
  if (rc) return rc;
  for (int i = 0; i < 36; i++) c_tangent[i] = res.c_tangent[i];
  return 0;
}

// micro/src/homogenize.c:51-57
extern "C" int homog_calculate_c_tangent_around_zero(double *c_tangent)
{
  double zero[MM_NVOI] = {0, 0, 0, 0, 0, 0};
  return homog_calculate_c_tangent(zero, c_tangent);
}

// micro/src/homogenize.c:35-49
extern "C" int homog_get_c_tangent(double *strain_mac, double **c_tangent)
{
  if (!g.init) return MM_ERR_STATE;
  if (!c_tangent) return MM_ERR_ARG;
  if (g.linear_materials) {
    *c_tangent = g.c_tangent_linear;
    return 0;
  }
  int rc = homog_calculate_c_tangent(strain_mac, g.c_tangent);
  *c_tangent = g.c_tangent;
  return rc;
}

// micro/src/main.c:116-151: one coupling request
extern "C" int micro_serve_message(micro_message *msg)
{
  if (!msg) return MM_ERR_ARG;
  switch (msg->action) {
    case ACTION_MICRO_CALC_STRESS: {
      double strain_ave[MM_NVOI];
      return homog_get_strain_stress(msg->strain_mac, strain_ave, msg->stress_ave);
    }
    case ACTION_MICRO_CALC_C_TANGENT: {
      double *c = nullptr;
      int rc = homog_get_c_tangent(msg->strain_mac, &c);
      if (rc) return rc;
      for (int i = 0; i < MM_NVOI * MM_NVOI; i++) msg->c_tangent_ave[i] = c[i];
      return 0;
    }
    case ACTION_MICRO_CALC_RHO:
      if (!g.init) return MM_ERR_STATE;
      msg->rho = g.cfg.rho;
      return 0;
    case ACTION_MICRO_END:
      return 0;
    default:
      return 1;
  }
}

extern "C" int micro_get_stats(micro_stats *st)
{
  if (!st) return MM_ERR_ARG;
  *st = g.st;
  return 0;
}

extern "C" int micro_get_sizes(int *nrow, int *nnz, int *nelm)
{
  if (!g.init) return MM_ERR_STATE;
  if (g.fe) {
    int r = 0, e = 0;
    mm::fe_sizes(g.fe, &r, &e, nullptr);
    if (nrow) *nrow = r;
    if (nnz) *nnz = 81;
    if (nelm) *nelm = e;
    return 0;
  }
  int r = 0, e = 0;
  int rc = mm_engine_rve_sizes(g.eng, kKey, &r, &e);
  if (rc) return rc;
  if (nrow) *nrow = r;
  if (nnz) *nnz = 81;
  if (nelm) *nelm = e;
  return 0;
}

extern "C" int micro_get_ell(int *cols, double *vals)
{
  if (!g.init) return MM_ERR_STATE;
  if (g.fe) return mm::fe_get_ell(g.fe, cols, vals);
  return mm_engine_download_ell(g.eng, kKey, cols, vals);
}

extern "C" int micro_get_elem_type(int *elem_type)
{
  if (!g.init) return MM_ERR_STATE;
  if (g.fe) return mm::fe_get_phase(g.fe, elem_type);
  return mm_engine_download_phase(g.eng, kKey, elem_type);
}

extern "C" int micro_get_displacement(double *x)
{
  if (!g.init) return MM_ERR_STATE;
  if (g.fe) return mm::fe_get_x(g.fe, x);
  return mm_engine_download_x(g.eng, kKey, 0, x);
}

// The element object for inspection calls: the element path's own, or a
// new one for the engine path's mesh (the caller destroys it then).
static int inspect_fe(mm::Fe **fe, bool *owned)
{
  *owned = false;
  if (g.fe) {
    *fe = g.fe;
    return 0;
  }
  int rc = fe_init(&g.cfg);
  if (rc) return rc;
  *fe = g.fe;
  g.fe = nullptr;
  *owned = true;
  return 0;
}

// assembly_res_ell (micro/src/assembly.c:33-82) for a given displacement:
// r = sum_e sum_gp B^T sigma w, boundary dofs zeroed, and |r|.  With
// MICRO_BC_PER_LM, x is the periodic fluctuation and `strain` the macro
// strain of the affine part.
extern "C" int micro_residual(const double *strain, const double *x, double *r, double *norm)
{
  if (!g.init) return MM_ERR_STATE;
  if (!x || !r) return MM_ERR_ARG;
  double zero[MM_NVOI] = {0, 0, 0, 0, 0, 0};
  mm::Fe *fe = nullptr;
  bool owned = false;
  int rc = inspect_fe(&fe, &owned);
  if (rc) return rc;
  double nrm = 0.0;
  if (!(rc = mm::fe_set_x(fe, x)) && !(rc = mm::fe_residual(fe, strain ? strain : zero, nullptr, &nrm)))
    rc = mm::fe_get_residual(fe, r);
  if (owned) mm::fe_destroy(fe);
  if (!rc && norm) *norm = nrm;
  return rc;
}

// micro_pvtu (micro/src/output.c:43-188): the .pvtu index and the one piece,
// ASCII, fields in the reference's order.  The displacement of the engine
// path is loaded into an element-assembly object to form the residual and
// the element strains and stresses on the GPU.
extern "C" int micro_write_vtu(const char *name)
{
  if (!g.init) return MM_ERR_STATE;
  if (!name || strlen(name) > 900) return MM_ERR_ARG;
  if (g.last_strain.size() != MM_NVOI) return MM_ERR_STATE;  // no homog_fe2 yet
  int rc = 0;
  mm::Fe *fe = nullptr;
  bool owned = false;
  std::vector<double> x;
  if ((rc = inspect_fe(&fe, &owned))) return rc;
  if (owned) {  // engine path: same mesh, USTRAIN, the engine's displacement
    int nrow = 0;
    mm::fe_sizes(fe, &nrow, nullptr, nullptr);
    x.resize(nrow);
    if ((rc = mm_engine_download_x(g.eng, kKey, 0, x.data())) || (rc = mm::fe_set_x(fe, x.data()))) {
      mm::fe_destroy(fe);
      return rc;
    }
  }
  const mm_rve_spec &rv = g.cfg.rve;
  const int nx = rv.nex + 1, ny = rv.ney + 1, nz = rv.nez + 1, nn = nx * ny * nz, nelm = rv.nex * rv.ney * rv.nez;
  int nrow = 0;
  mm::fe_sizes(fe, &nrow, nullptr, nullptr);
  std::vector<double> xu(nrow), res(nrow), es((size_t)nelm * 6), ss((size_t)nelm * 6);
  std::vector<int> et(nelm);
  double norm = 0.0, sa[6], sv[6];
  const double *strain = g.last_strain.data();
  if ((rc = mm::fe_get_x(fe, xu.data())) || (rc = mm::fe_residual(fe, strain, nullptr, &norm)) ||
      (rc = mm::fe_get_residual(fe, res.data())) || (rc = mm::fe_averages(fe, strain, sa, sv, es.data(), ss.data())) ||
      (rc = mm::fe_get_phase(fe, et.data()))) {
    if (owned) mm::fe_destroy(fe);
    return rc;
  }
  if (owned) mm::fe_destroy(fe);
  const bool per = g.cfg.fe2_bc == MICRO_BC_PER_LM;
  const double hx = rv.lx / rv.nex, hy = rv.ly / rv.ney, hz = rv.lz / rv.nez;
  // per structured node: the displacement (E X + fluctuation when periodic)
  // and the residual of its unknown
  auto unk = [&](int ix, int iy, int iz) {
    return per ? (ix % rv.nex) + (iy % rv.ney) * rv.nex + (iz % rv.nez) * rv.nex * rv.ney : ix + iy * nx + iz * nx * ny;
  };
  char fname[1024];
  snprintf(fname, sizeof(fname), "%s.pvtu", name);
  FILE *fm = fopen(fname, "w");
  if (!fm) return MM_ERR_ARG;
  fprintf(fm,
          "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
          "<PUnstructuredGrid GhostLevel=\"0\">\n<PPoints>\n"
          "<PDataArray type=\"Float32\" Name=\"Position\" NumberOfComponents=\"3\"/>\n</PPoints>\n<PCells>\n"
          "<PDataArray type=\"Int32\" Name=\"connectivity\" NumberOfComponents=\"1\"/>\n"
          "<PDataArray type=\"Int32\" Name=\"offsets\"      NumberOfComponents=\"1\"/>\n"
          "<PDataArray type=\"UInt8\" Name=\"types\"        NumberOfComponents=\"1\"/>\n</PCells>\n"
          "<PPointData Vectors=\"displ\">\n"
          "<PDataArray type=\"Float64\" Name=\"displ\"    NumberOfComponents=\"3\" />\n"
          "<PDataArray type=\"Float64\" Name=\"residual\" NumberOfComponents=\"3\" />\n</PPointData>\n"
          "<PCellData>\n<PDataArray type=\"Int32\"   Name=\"part\"   NumberOfComponents=\"1\"/>\n"
          "<PDataArray type=\"Float64\" Name=\"strain\" NumberOfComponents=\"6\"/>\n"
          "<PDataArray type=\"Float64\" Name=\"stress\" NumberOfComponents=\"6\"/>\n"
          "<PDataArray type=\"Int32\"   Name=\"elem_type\" NumberOfComponents=\"1\"/>\n</PCellData>\n"
          "<Piece Source=\"%s_0.vtu\"/>\n</PUnstructuredGrid>\n</VTKFile>\n",
          name);
  fclose(fm);
  snprintf(fname, sizeof(fname), "%s_0.vtu", name);
  fm = fopen(fname, "w");
  if (!fm) return MM_ERR_ARG;
  fprintf(fm, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
              "<UnstructuredGrid>\n<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n<Points>\n"
              "<DataArray type=\"Float32\" Name=\"Position\" NumberOfComponents=\"3\" format=\"ascii\">\n",
          nn, nelm);
  for (int n = 0; n < nn; n++)
    fprintf(fm, "%e %e %e\n", (n % nx) * hx, ((n / nx) % ny) * hy, (n / (nx * ny)) * hz);
  fprintf(fm, "</DataArray>\n</Points>\n<Cells>\n"
              "<DataArray type=\"Int32\" Name=\"connectivity\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) {
    const int n0 = e % rv.nex + ((e / rv.nex) % rv.ney) * nx + (e / (rv.nex * rv.ney)) * nx * ny;
    fprintf(fm, "%d %d %d %d %d %d %d %d\n", n0, n0 + 1, n0 + nx + 1, n0 + nx, n0 + nx * ny, n0 + nx * ny + 1,
            n0 + nx * ny + nx + 1, n0 + nx * ny + nx);
  }
  fprintf(fm, "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) fprintf(fm, "%d ", 8 * (e + 1));
  fprintf(fm, "\n</DataArray>\n<DataArray type=\"UInt8\"  Name=\"types\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) fprintf(fm, "12 ");
  fprintf(fm, "\n</DataArray>\n</Cells>\n<PointData Vectors=\"displ\">\n"
              "<DataArray type=\"Float64\" Name=\"displ\" NumberOfComponents=\"3\" format=\"ascii\" >\n");
  for (int n = 0; n < nn; n++) {
    const int ix = n % nx, iy = (n / nx) % ny, iz = n / (nx * ny), u = unk(ix, iy, iz);
    const double X[3] = {ix * hx, iy * hy, iz * hz};
    double a[3] = {0.0, 0.0, 0.0};
    if (per) {
      a[0] = strain[0] * X[0] + strain[3] / 2 * X[1] + strain[5] / 2 * X[2];
      a[1] = strain[3] / 2 * X[0] + strain[1] * X[1] + strain[4] / 2 * X[2];
      a[2] = strain[5] / 2 * X[0] + strain[4] / 2 * X[1] + strain[2] * X[2];
    }
    fprintf(fm, "%e %e %e\n", xu[3 * u] + a[0], xu[3 * u + 1] + a[1], xu[3 * u + 2] + a[2]);
  }
  fprintf(fm, "</DataArray>\n<DataArray type=\"Float64\" Name=\"residual\" NumberOfComponents=\"3\" format=\"ascii\" >\n");
  for (int n = 0; n < nn; n++) {
    const int u = unk(n % nx, (n / nx) % ny, n / (nx * ny));
    fprintf(fm, "%e %e %e\n", res[3 * u], res[3 * u + 1], res[3 * u + 2]);
  }
  fprintf(fm, "</DataArray>\n</PointData>\n<CellData>\n"
              "<DataArray type=\"Int32\" Name=\"part\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) fprintf(fm, "0 ");
  fprintf(fm, "\n</DataArray>\n<DataArray type=\"Float64\" Name=\"strain\" NumberOfComponents=\"6\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) {
    for (int v = 0; v < 6; v++) fprintf(fm, "%lf ", es[(size_t)e * 6 + v]);
    fprintf(fm, "\n");
  }
  fprintf(fm, "</DataArray>\n<DataArray type=\"Float64\" Name=\"stress\" NumberOfComponents=\"6\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) {
    for (int v = 0; v < 6; v++) fprintf(fm, "%lf ", ss[(size_t)e * 6 + v]);
    fprintf(fm, "\n");
  }
  fprintf(fm, "</DataArray>\n<DataArray type=\"Int32\" Name=\"elem_type\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < nelm; e++) fprintf(fm, "%d ", et[e]);
  fprintf(fm, "\n</DataArray>\n</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  fclose(fm);
  return 0;
}
