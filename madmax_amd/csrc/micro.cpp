// micro.cpp -- the reference's process-global micro homogenization entry
// points (micro/inc/micro.h:150-157, micro/src/homogenize.c) served by a
// one-RVE engine.  The RVE stays resident in HBM between calls (assembled
// once); every solve starts from zero interior displacement.
#include <chrono>
#include <cstdio>
#include <cstring>

#include "micro_api.h"

namespace {

constexpr uint64_t kKey = 1;  // resident RVE key inside the private engine

struct MicroGlobal {
  int init = 0;
  micro_config cfg;
  mm_engine *eng = nullptr;
  double c_tangent_linear[36];  // micro/inc/micro.h:90
  double c_tangent[36];         // micro/inc/micro.h:91
  bool linear_materials = true;      // every material here is MAT_ELASTIC
  bool c_linear_calculated = false;  // micro/inc/micro.h:101
  micro_stats st;
  uint64_t next_id = 1;
} g;

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
  return 0;
}

extern "C" int micro_finalize(void)
{
  if (g.eng) mm_engine_destroy(g.eng);
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
  g.cfg = *cfg;
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
  return mm_engine_download_ell(g.eng, kKey, cols, vals);
}

extern "C" int micro_get_elem_type(int *elem_type)
{
  if (!g.init) return MM_ERR_STATE;
  return mm_engine_download_phase(g.eng, kKey, elem_type);
}

extern "C" int micro_get_displacement(double *x)
{
  if (!g.init) return MM_ERR_STATE;
  return mm_engine_download_x(g.eng, kKey, 0, x);
}
