// engine.hip -- host side of the RVE engine: HBM arena, admission by
// footprint, resident-RVE cache, hipGraph-captured solver steps, retirement,
// and the serving worker thread.  C-ABI in include/micro_api.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "micro_api.h"
#include "mm_internal.h"

using namespace mm;

namespace {

// First-fit allocator of index ranges [0, cap).
struct RangeAlloc {
  std::map<long long, long long> free_;  // start -> length
  long long cap = 0;
  void init(long long c)
  {
    free_.clear();
    cap = c;
    if (c > 0) free_[0] = c;
  }
  long long alloc(long long n)
  {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const long long s = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[s + n] = len - n;
        return s;
      }
    }
    return -1;
  }
  void release(long long s, long long n)
  {
    auto it = free_.emplace(s, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  long long high_water() const
  {
    if (free_.empty()) return cap;
    auto last = std::prev(free_.end());
    return (last->first + last->second == cap) ? last->first : cap;
  }
};

enum SlotState { SL_FREE = 0, SL_ACTIVE = 1, SL_CACHED = 2 };

struct HostSlot {
  int state = SL_FREE;
  mm_request req;
  // load cases of the request that last left displacements in the slot
  double prev_load[MM_MAX_RHS][MM_NVOI];
  int prev_nload = 0;
  long long first_slab = 0, nslab = 0, elem_base = 0, nelm = 0;
  int nrow = 0;
  double t_submit = 0, t_admit = 0;
  uint64_t lru = 0;
};

struct Queued {
  mm_request req;
  double t_submit;
};

bool same_geometry(const mm_rve_spec &a, const mm_rve_spec &b)
{
  if (a.nex != b.nex || a.ney != b.ney || a.nez != b.nez || a.lx != b.lx || a.ly != b.ly ||
      a.lz != b.lz || a.nfib != b.nfib)
    return false;
  for (int i = 0; i < a.nfib; i++)
    if (a.fib_x[i] != b.fib_x[i] || a.fib_y[i] != b.fib_y[i] || a.fib_r[i] != b.fib_r[i]) return false;
  for (int i = 0; i < 2; i++)
    if (a.young[i] != b.young[i] || a.poisson[i] != b.poisson[i]) return false;
  return true;
}

int spec_valid(const mm_rve_spec &s)
{
  if (s.nex < 1 || s.ney < 1 || s.nez < 1) return 0;
  if (!(s.lx > 0) || !(s.ly > 0) || !(s.lz > 0)) return 0;
  if (s.nfib < 0 || s.nfib > MM_MAX_FIBERS) return 0;
  const double nn = (double)(s.nex + 1) * (s.ney + 1) * (s.nez + 1);
  if (nn * 3 > 2.0e9) return 0;
  for (int i = 0; i < 2; i++) {
    if (!std::isfinite(s.young[i]) || !(s.young[i] > 0)) return 0;
    if (!(s.poisson[i] > -1.0 && s.poisson[i] < 0.5)) return 0;
  }
  for (int i = 0; i < s.nfib; i++)
    if (!std::isfinite(s.fib_x[i]) || !std::isfinite(s.fib_y[i]) || !std::isfinite(s.fib_r[i]))
      return 0;
  return 1;
}

template <typename T>
int ensure_pinned(T **p, size_t *cap, size_t n)
{
  if (n <= *cap) return 0;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  size_t c = std::max<size_t>(n, 2 * *cap);
  if (hipHostMalloc((void **)p, c * sizeof(T), hipHostMallocDefault) != hipSuccess) {
    *cap = 0;
    return MM_ERR_HIP;
  }
  *cap = c;
  return 0;
}

template <typename T>
int ensure_device(T **p, size_t *cap, size_t n)
{
  if (n <= *cap) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  size_t c = std::max<size_t>(n, 2 * *cap);
  if (hipMalloc((void **)p, c * sizeof(T)) != hipSuccess) {
    *cap = 0;
    return MM_ERR_HIP;
  }
  *cap = c;
  return 0;
}

}  // namespace

struct mm_engine {
  mm_engine_config cfg;
  int SN = 128, K = 1;
  bool fp64 = false;
  bool dict_mode = false;  // MM_ELL_DICT: no per-row ELL arrays, dictionary + pattern bytes
  size_t tsz = 4;
  hipStream_t stream = nullptr;
  // solver_streams == 2: the resident RVEs are split into two groups of
  // contiguous slabs, each stepped on its own stream (stream, stream2), so
  // that one group's kernels fill the other's launch gaps and tails.
  // ev_fork / ev_join order stream2 after the admission work and the engine
  // stream after stream2's steps.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // arena
  long long nslab_cap = 0, elem_cap = 0;
  int rve_cap = 0;
  int *bcols = nullptr;
  void *vals = nullptr, *dinv = nullptr;
  double *x64 = nullptr;
  void *d = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *part = nullptr;
  int *slab_rve = nullptr, *slab_node0 = nullptr;
  RveDev *rves = nullptr;
  uint8_t *elem_mat = nullptr;
  uint8_t *npat = nullptr;
  void *dict = nullptr;      // solver dictionary (TS); aliases dict64 in FP64 mode
  double *dict64 = nullptr;  // FP64 dictionary, the residual's matrix
  double *ke_tab = nullptr, *c_tab = nullptr, *bw_tab = nullptr;
  RveOut *out_host = nullptr, *out_dev = nullptr;
  size_t arena_bytes = 0;
  // staging
  RveDev *st_states = nullptr;
  int *st_slots = nullptr, *st_ke = nullptr, *st_slabs = nullptr;
  int2 *st_phase = nullptr;
  size_t cap_states = 0, cap_slots = 0, cap_ke = 0, cap_slabs = 0, cap_phase = 0;
  RveDev *dv_states = nullptr;
  int *dv_slots = nullptr, *dv_ke = nullptr, *dv_slabs = nullptr;
  int2 *dv_phase = nullptr;
  size_t dcap_states = 0, dcap_slots = 0, dcap_ke = 0, dcap_slabs = 0, dcap_phase = 0;
  // bookkeeping
  RangeAlloc slabs, elems;
  std::vector<HostSlot> slots;
  std::vector<int> free_slots;
  std::unordered_map<uint64_t, int> warm_map;
  uint64_t lru_clock = 0;
  int n_active = 0;
  // queue / results
  std::mutex qmu;
  std::condition_variable qcv;
  std::deque<Queued> queue;
  std::deque<mm_result> results;
  // solver launches run over the slabs of the RVEs being solved, one list
  // per stream group g at act_list + g * (nslab_cap + 1): [0] is the count,
  // [1..] the slabs in ascending order (device), staged in st_act (pinned);
  // rebuilt when an RVE is admitted or retired.  With the plane march
  // (march_planes > 0) kernel A runs over the tiles of the RVEs that march,
  // listed by their slabs at act_list + (2 + g) * (nslab_cap + 1), with
  // a_lds[g] bytes of dynamic LDS; a_slab[g]: the group also holds RVEs
  // that go slab by slab; s_big[g]: it holds an RVE of kBigSlabs+ slabs
  // (kernel S runs after A).
  int *act_list = nullptr, *st_act = nullptr;
  bool act_dirty = true;
  long long act_n[2] = {0, 0}, tile_n[2] = {0, 0};
  int a_lds[2] = {0, 0}, a_slab[2] = {0, 0}, s_big[2] = {0, 0};
  // one captured graph per set of launch shapes (grids: powers of two >= 64
  // covering each group's slabs and tiles, 0 for an empty group; kernel A's
  // LDS); the kernels read the active count from their list
  std::map<std::array<int, 10>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  // stats / threading
  std::mutex step_mu;
  long long completed = 0, nsteps = 0;
  double spmv_ms = 0.0;
  std::thread worker;
  std::atomic<bool> running{false};
  std::chrono::steady_clock::time_point t0;
  int last_error = 0;
  // profile_step_kernel: event pairs around the kernel A launches queued in
  // the current step, and the durations recorded so far
  std::vector<hipEvent_t> prof_ev;
  int prof_pending = 0;
  std::vector<double> prof_ms;

  EngPtrs ptrs() const
  {
    EngPtrs P;
    P.bcols = bcols;
    P.vals = vals;
    P.dinv = dinv;
    P.x64 = x64;
    P.d = d;
    P.z = z;
    P.p = p;
    P.q = q;
    P.part = part;
    P.part_stride = nslab_cap;
    P.slab_rve = slab_rve;
    P.slab_node0 = slab_node0;
    P.rves = rves;
    P.out = out_dev;
    P.ke_tab = ke_tab;
    P.c_tab = c_tab;
    P.bw_tab = bw_tab;
    P.elem_mat = elem_mat;
    P.npat = npat;
    P.act_list = act_list;  // group 0; run_steps points group 1 at its list
    P.act_tile = act_list + 2 * ((size_t)nslab_cap + 1);
    P.a_grid = 0;
    P.a_lds = 0;
    P.a_slab = 0;
    P.s_big = 0;
    P.dict = dict;
    P.dict64 = dict64;
    P.SN = SN;
    P.sn_shift = __builtin_ctz((unsigned)SN);
    return P;
  }
  SolveParams params() const
  {
    SolveParams sp;
    sp.nrhs = K;
    sp.slab_nodes = SN;
    sp.rtol = cfg.rtol;
    sp.abs_tol = cfg.abs_tol;
    sp.inner_rtol = cfg.inner_rtol;
    sp.max_inner_its = cfg.max_inner_its;
    sp.max_sweeps = cfg.max_sweeps;
    return sp;
  }
  double now() const
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  int rows_per_slab() const { return 3 * SN; }
  size_t bytes_per_slab() const
  {
    const size_t RS = (size_t)rows_per_slab();
    const size_t ell = dict_mode ? 0 : RS * kW * tsz + (size_t)SN * 27 * 4;
    return ell + RS * (K * (8 + 4 * tsz) + tsz) + kPartDoubles * 8 + 8 + (size_t)SN * 2 + 1;
  }
};

static void destroy_streams(mm_engine *e)
{
  for (hipEvent_t ev : e->prof_ev) (void)hipEventDestroy(ev);
  e->prof_ev.clear();
  if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
  if (e->ev_join) (void)hipEventDestroy(e->ev_join);
  if (e->stream2) (void)hipStreamDestroy(e->stream2);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->ev_fork = e->ev_join = nullptr;
  e->stream2 = e->stream = nullptr;
}

static int engine_free_arena(mm_engine *e)
{
  void *ptrs[] = {e->bcols, e->vals, e->dinv, e->x64, e->d, e->z, e->p, e->q, e->part,
                  e->slab_rve, e->slab_node0, e->rves, e->elem_mat, e->npat,
                  (e->dict != (void *)e->dict64) ? e->dict : nullptr, e->dict64, e->ke_tab, e->c_tab,
                  e->bw_tab,
                  e->dv_states, e->dv_slots, e->dv_ke, e->dv_slabs, e->dv_phase};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (e->act_list) (void)hipFree(e->act_list);
  void *hp[] = {e->out_host, e->st_states, e->st_slots, e->st_ke, e->st_slabs, e->st_phase, e->st_act};
  for (void *p : hp)
    if (p) (void)hipHostFree(p);
  return 0;
}

extern "C" int mm_engine_config_default(mm_engine_config *c)
{
  if (!c) return MM_ERR_ARG;
  memset(c, 0, sizeof(*c));
  c->device = 0;
  c->hbm_budget = 0;
  c->max_rves = 0;
  c->slab_nodes = 0;
  c->nrhs = 6;
  c->precision = MM_PREC_MIXED;
  c->rtol = 1.0e-10;
  c->abs_tol = 0.0;
  c->inner_rtol = 1.0e-5;
  c->max_inner_its = 20000;
  c->max_sweeps = 8;
  c->steps_per_launch = 16;
  c->use_graph = 1;
  c->sync_every_step = 0;
  c->ell_storage = MM_ELL_DICT;
  c->initial_guess = MM_GUESS_AFFINE;
  c->solver_streams = 2;
  c->profile_step_kernel = 0;
  c->march_planes = 8;
  c->big_slabs = kBigSlabs;
  return 0;
}

extern "C" int mm_engine_create(const mm_engine_config *cfg, mm_engine **out)
{
  if (!cfg || !out) return MM_ERR_ARG;
  if (!eng_supported_rhs(cfg->nrhs)) {
    fprintf(stderr, "madmax_amd: nrhs %d not supported (1, 6 or 7)\n", cfg->nrhs);
    return MM_ERR_ARG;
  }
  if (cfg->slab_nodes != 0 && cfg->slab_nodes != 64 && cfg->slab_nodes != 128 && cfg->slab_nodes != 256)
    return MM_ERR_ARG;
  if (cfg->precision != MM_PREC_FP64 && cfg->precision != MM_PREC_MIXED) return MM_ERR_ARG;
  if (cfg->ell_storage != MM_ELL_FULL && cfg->ell_storage != MM_ELL_DICT) return MM_ERR_ARG;
  if (cfg->initial_guess != MM_GUESS_ZERO && cfg->initial_guess != MM_GUESS_AFFINE) return MM_ERR_ARG;
  if (cfg->solver_streams != 1 && cfg->solver_streams != 2) return MM_ERR_ARG;
  if (cfg->march_planes < 0 || cfg->march_planes > 4096) return MM_ERR_ARG;
  if (cfg->big_slabs < 1) return MM_ERR_ARG;
  if (!(cfg->rtol >= 0) || !(cfg->inner_rtol > 0) || cfg->max_inner_its < 1 || cfg->max_sweeps < 1)
    return MM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    fprintf(stderr, "madmax_amd: no HIP device available\n");
    return MM_ERR_HIP;
  }
  if (cfg->device < 0 || cfg->device >= ndev) return MM_ERR_ARG;
  MM_HIP(hipSetDevice(cfg->device));
  mm_engine *e = new mm_engine();
  e->cfg = *cfg;
  if (e->cfg.steps_per_launch < 1) e->cfg.steps_per_launch = 1;
  if (e->cfg.sync_every_step) {
    e->cfg.steps_per_launch = 1;
    e->cfg.use_graph = 0;
  }
  if (e->cfg.profile_step_kernel) {  // every kernel A launch bracketed by events
    e->cfg.use_graph = 0;
    e->cfg.solver_streams = 1;
  }
  e->SN = cfg->slab_nodes ? cfg->slab_nodes : 128;
  e->K = cfg->nrhs;
  e->fp64 = (cfg->precision == MM_PREC_FP64);
  e->tsz = e->fp64 ? 8 : 4;
  e->dict_mode = (cfg->ell_storage == MM_ELL_DICT);
  e->rve_cap = cfg->max_rves > 0 ? cfg->max_rves : 4096;
  e->t0 = std::chrono::steady_clock::now();
  int rc = 0;
  size_t budget = cfg->hbm_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete e;
      return MM_ERR_HIP;
    }
    budget = (size_t)(0.8 * (double)fr);
  }
  const size_t per_rve = (size_t)kNmat * (kKe + 36) * 8 + 144 * 8 + sizeof(RveDev) + sizeof(RveOut) +
                         (size_t)kDictPat * kDictEntry * 8 +
                         ((e->dict_mode && !e->fp64) ? (size_t)kDictPat * kDictEntry * 4 : 0);
  if (budget <= (size_t)e->rve_cap * per_rve + e->bytes_per_slab()) {
    delete e;
    return MM_ERR_ALLOC;
  }
  e->nslab_cap = (long long)((budget - (size_t)e->rve_cap * per_rve) / e->bytes_per_slab());
  e->nslab_cap -= e->nslab_cap % 8;  // solver grids are multiples of 8 (XCD-aware order)
  if (e->nslab_cap < 8) {
    delete e;
    return MM_ERR_ALLOC;
  }
  e->elem_cap = e->nslab_cap * e->SN + 64;
  const size_t nrows = (size_t)e->nslab_cap * e->rows_per_slab();
  const size_t slots = nrows * kW;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return MM_ERR_HIP;
  }
  if (e->cfg.solver_streams == 2 &&
      (hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess)) {
    destroy_streams(e);
    delete e;
    return MM_ERR_HIP;
  }
#define EALLOC(ptr, bytes)                                                  \
  do {                                                                      \
    if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {                \
      fprintf(stderr, "madmax_amd: arena allocation of %zu bytes failed\n", \
              (size_t)(bytes));                                             \
      rc = MM_ERR_ALLOC;                                                    \
      goto fail;                                                            \
    }                                                                       \
    e->arena_bytes += (bytes);                                              \
  } while (0)
  if (!e->dict_mode) {
    EALLOC(e->bcols, (size_t)e->nslab_cap * e->SN * 27 * 4);
    EALLOC(e->vals, slots * e->tsz);
  }
  EALLOC(e->dict64, (size_t)e->rve_cap * kDictPat * kDictEntry * 8);
  if (e->dict_mode) {
    if (e->fp64) e->dict = e->dict64;
    else EALLOC(e->dict, (size_t)e->rve_cap * kDictPat * kDictEntry * 4);
  }
  EALLOC(e->npat, (size_t)e->nslab_cap * e->SN);
  EALLOC(e->act_list, 4 * ((size_t)e->nslab_cap + 1) * 4);
  EALLOC(e->dinv, nrows * e->tsz);
  EALLOC(e->x64, nrows * e->K * 8);
  EALLOC(e->d, nrows * e->K * e->tsz);
  EALLOC(e->z, nrows * e->K * e->tsz);
  EALLOC(e->p, nrows * e->K * e->tsz);
  EALLOC(e->q, nrows * e->K * e->tsz);
  EALLOC(e->part, (size_t)e->nslab_cap * kPartDoubles * 8);
  EALLOC(e->slab_rve, (size_t)e->nslab_cap * 4);
  EALLOC(e->slab_node0, (size_t)e->nslab_cap * 4);
  EALLOC(e->rves, (size_t)e->rve_cap * sizeof(RveDev));
  EALLOC(e->elem_mat, (size_t)e->elem_cap);
  EALLOC(e->ke_tab, (size_t)e->rve_cap * kNmat * kKe * 8);
  EALLOC(e->c_tab, (size_t)e->rve_cap * kNmat * 36 * 8);
  EALLOC(e->bw_tab, (size_t)e->rve_cap * 144 * 8);
#undef EALLOC
  if (hipHostMalloc((void **)&e->out_host, (size_t)e->rve_cap * sizeof(RveOut), hipHostMallocMapped) !=
      hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  memset(e->out_host, 0, (size_t)e->rve_cap * sizeof(RveOut));
  if (hipHostMalloc((void **)&e->st_act, 4 * ((size_t)e->nslab_cap + 1) * 4, hipHostMallocDefault) !=
      hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  if (hipHostGetDevicePointer((void **)&e->out_dev, e->out_host, 0) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  if (hipMemsetAsync(e->slab_rve, 0xff, (size_t)e->nslab_cap * 4, e->stream) != hipSuccess ||
      hipMemsetAsync(e->rves, 0, (size_t)e->rve_cap * sizeof(RveDev), e->stream) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto fail;
  }
  rc = upload_slot_tables(e->stream);
  if (rc) goto fail;
  e->slabs.init(e->nslab_cap);
  e->elems.init(e->elem_cap);
  e->slots.resize(e->rve_cap);
  for (int i = e->rve_cap - 1; i >= 0; i--) e->free_slots.push_back(i);
  *out = e;
  return 0;
fail:
  engine_free_arena(e);
  destroy_streams(e);
  delete e;
  return rc;
}

extern "C" int mm_engine_stop(mm_engine *e);

extern "C" int mm_engine_destroy(mm_engine *e)
{
  if (!e) return MM_ERR_ARG;
  mm_engine_stop(e);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (auto &g : e->graphs) {
    (void)hipGraphExecDestroy(g.second.second);
    (void)hipGraphDestroy(g.second.first);
  }
  engine_free_arena(e);
  destroy_streams(e);
  delete e;
  return 0;
}

static void footprint(const mm_engine *e, const mm_rve_spec &s, int *nrow, long long *nslab,
                      long long *nelm)
{
  const long long nn = (long long)(s.nex + 1) * (s.ney + 1) * (s.nez + 1);
  *nrow = (int)(3 * nn);
  *nslab = (nn + e->SN - 1) / e->SN;
  *nelm = (long long)s.nex * s.ney * s.nez;
}

// The tangent of a block is formed column by column from the load cases
// (finish_rve): UNIT divides the stress of load i by its strain component i,
// FD divides the difference of loads i+1 and 0 by their difference in
// component i (micro/src/homogenize.c:59-94).  Both are right only for loads
// of exactly that shape, so any other load is refused here instead of
// returning a wrong tangent with status MM_OK.
extern "C" int mm_request_validate(const mm_request *req, int nrhs)
{
  if (!req) return MM_ERR_ARG;
  if (nrhs < 1 || nrhs > MM_MAX_RHS) return MM_ERR_ARG;
  if (!spec_valid(req->rve) || req->nload < 1 || req->nload > nrhs) return MM_ERR_ARG;
  if (req->tangent_mode != MM_TANGENT_NONE && req->tangent_mode != MM_TANGENT_UNIT &&
      req->tangent_mode != MM_TANGENT_FD)
    return MM_ERR_ARG;
  if (req->tangent_mode == MM_TANGENT_UNIT && req->nload < 6) return MM_ERR_ARG;
  if (req->tangent_mode == MM_TANGENT_FD && req->nload < 7) return MM_ERR_ARG;
  for (int k = 0; k < req->nload; k++)
    for (int i = 0; i < MM_NVOI; i++)
      if (!std::isfinite(req->strain[k][i])) return MM_ERR_ARG;
  if (req->tangent_mode == MM_TANGENT_UNIT)
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < MM_NVOI; j++)
        if ((j == i) != (req->strain[i][j] != 0.0)) return MM_ERR_ARG;  // load i must be a e_i, a != 0
  if (req->tangent_mode == MM_TANGENT_FD)
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < MM_NVOI; j++)
        if ((j == i) != (req->strain[i + 1][j] != req->strain[0][j])) return MM_ERR_ARG;  // eps + d e_i
  if (req->warm_start != 0 && req->warm_start != 1) return MM_ERR_ARG;
  return MM_OK;
}

extern "C" int mm_engine_submit(mm_engine *e, const mm_request *req)
{
  if (!e || !req) return MM_ERR_ARG;
  const int vrc = mm_request_validate(req, e->K);
  if (vrc) return vrc;
  int nrow;
  long long nslab, nelm;
  footprint(e, req->rve, &nrow, &nslab, &nelm);
  if (nslab > e->nslab_cap || nelm > e->elem_cap) return MM_REJECTED;
  // the SpMV addresses an RVE's run of a solver vector with 32-bit byte
  // offsets (buffer loads), so one run must stay below 2^31 bytes
  if ((long long)nslab * e->K * e->rows_per_slab() * (long long)e->tsz > INT_MAX) return MM_REJECTED;
  {
    std::lock_guard<std::mutex> lk(e->qmu);
    e->queue.push_back(Queued{*req, e->now()});
  }
  e->qcv.notify_all();
  return 0;
}

static void release_slot_resources(mm_engine *e, int slot)
{
  HostSlot &h = e->slots[slot];
  e->slabs.release(h.first_slab, h.nslab);
  e->elems.release(h.elem_base, h.nelm);
  (void)hipMemsetAsync(e->slab_rve + h.first_slab, 0xff, (size_t)h.nslab * 4, e->stream);
  if (h.req.warm_key) {
    auto it = e->warm_map.find(h.req.warm_key);
    if (it != e->warm_map.end() && it->second == slot) e->warm_map.erase(it);
  }
  h.state = SL_FREE;
  e->free_slots.push_back(slot);
}

static bool evict_one(mm_engine *e)
{
  int best = -1;
  for (int i = 0; i < e->rve_cap; i++)
    if (e->slots[i].state == SL_CACHED && (best < 0 || e->slots[i].lru < e->slots[best].lru)) best = i;
  if (best < 0) return false;
  release_slot_resources(e, best);
  return true;
}

// Warm-start coefficients.  The kept displacements x_old[j] solve the loads
// old[j]; for linear materials the solution is linear in the load, so when
// six of the old loads span the strain space, a new load e = sum_j w_j old[j]
// starts from sum_j w_j x_old[j], which is already converged up to the old
// solves' tolerance.  Otherwise load k starts from x_old[k] plus the affine
// field of the strain increment e_k - old[k].  W[k][j] are the weights and
// D[k] the strain whose affine field is added (zero up to rounding in the
// first case).
static void warm_coefficients(const double (*old)[MM_NVOI], int nold, const double (*nw)[MM_NVOI], int nnew,
                              double (*W)[MM_MAX_RHS], double (*D)[MM_NVOI])
{
  for (int k = 0; k < MM_MAX_RHS; k++) {
    for (int j = 0; j < MM_MAX_RHS; j++) W[k][j] = 0.0;
    for (int i = 0; i < MM_NVOI; i++) D[k][i] = 0.0;
  }
  // old loads that span the strain space: modified Gram-Schmidt, a load is
  // kept when its component outside the span of the kept ones is more than
  // 1e-6 of its length
  int pick[MM_NVOI], np = 0;
  double Q[MM_NVOI][MM_NVOI];
  for (int j = 0; j < nold && np < MM_NVOI; j++) {
    double v[MM_NVOI], len0 = 0.0, len = 0.0;
    for (int i = 0; i < MM_NVOI; i++) {
      v[i] = old[j][i];
      len0 += v[i] * v[i];
    }
    for (int q = 0; q < np; q++) {
      double dot = 0.0;
      for (int i = 0; i < MM_NVOI; i++) dot += Q[q][i] * v[i];
      for (int i = 0; i < MM_NVOI; i++) v[i] -= dot * Q[q][i];
    }
    for (int i = 0; i < MM_NVOI; i++) len += v[i] * v[i];
    if (len0 > 0.0 && len > 1e-12 * len0) {
      const double inv = 1.0 / std::sqrt(len);
      for (int i = 0; i < MM_NVOI; i++) Q[np][i] = v[i] * inv;
      pick[np++] = j;
    }
  }
  for (int k = 0; k < nnew; k++) {
    if (np == MM_NVOI) {
      // solve M w = e, column c of M the picked load pick[c] (partial pivoting)
      double A[MM_NVOI][MM_NVOI + 1];
      for (int r = 0; r < MM_NVOI; r++) {
        for (int c = 0; c < MM_NVOI; c++) A[r][c] = old[pick[c]][r];
        A[r][MM_NVOI] = nw[k][r];
      }
      for (int c = 0; c < MM_NVOI; c++) {
        int p = c;
        for (int r = c + 1; r < MM_NVOI; r++)
          if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (p != c)
          for (int cc = 0; cc <= MM_NVOI; cc++) std::swap(A[p][cc], A[c][cc]);
        for (int r = c + 1; r < MM_NVOI; r++) {
          const double f = A[r][c] / A[c][c];
          for (int cc = c; cc <= MM_NVOI; cc++) A[r][cc] -= f * A[c][cc];
        }
      }
      double w[MM_NVOI];
      for (int c = MM_NVOI - 1; c >= 0; c--) {
        double t = A[c][MM_NVOI];
        for (int cc = c + 1; cc < MM_NVOI; cc++) t -= A[c][cc] * w[cc];
        w[c] = t / A[c][c];
      }
      for (int c = 0; c < MM_NVOI; c++) W[k][pick[c]] = w[c];
      for (int i = 0; i < MM_NVOI; i++) {
        double s = nw[k][i];
        for (int c = 0; c < MM_NVOI; c++) s -= w[c] * old[pick[c]][i];
        D[k][i] = s;
      }
    } else {
      if (k < nold) W[k][k] = 1.0;
      for (int i = 0; i < MM_NVOI; i++) D[k][i] = nw[k][i] - (k < nold ? old[k][i] : 0.0);
    }
  }
}

// Plane-march tiling of an RVE of nx x ny x nz nodes in nslab slabs (kernel
// A's tiles, RveDev::mcpp ..): in-plane chunks of at most SN nodes times
// runs of at most march_planes z-planes (even splits) times groups of load
// cases (two in FP32, one in FP64, whose ring takes twice the bytes), the
// run doubled until the tiles fit the RVE's slabs.  mcpp = 0 (slab by slab)
// when the march is off, with one load case (27 gathers per node cost less
// than a tile's fixed work), or when the window (SN + 2 nx + 2 nodes) is
// longer than two positions per lane or the march's LDS exceeds its bound.  Fixed by
// the RVE and the engine configuration alone, so an RVE's bits do not depend
// on the RVEs solved beside it.
struct MarchTiling {
  int mcpp = 0, mL = 0, mnzc = 0, mnt = 0, mncg = 0;
  long long tiles() const { return (long long)mnt * mncg; }
};
static MarchTiling march_tiling(const mm_engine *e, const mm_rve_spec &s, long long nslab)
{
  MarchTiling m;
  const int nx = s.nex + 1, ny = s.ney + 1, nz = s.nez + 1, SN = e->SN;
  if (e->cfg.march_planes <= 0 || 2 * nx + 2 > SN) return m;
  if (e->K < 2 || march_lds_bytes(SN, nx, ny, e->K, (int)e->tsz, e->dict_mode) > kMarchMaxLds)
    return m;
  const long long cpp = ((long long)nx * ny + SN - 1) / SN;
  long long L = e->cfg.march_planes;
  // an RVE of 2048+ slabs can fill the GPU alone (a tile per 1024 slabs of
  // it keeps every CU busy): its runs shrink to the power of two nearest
  // below nslab / 1024 planes (at least 2)
  if (nslab >= 2048) {
    long long cap = 2;
    while (cap * 2 <= nslab / 1024) cap *= 2;
    L = std::min(L, cap);
  }
  const long long ncg = (e->K + march_cases((int)e->tsz) - 1) / march_cases((int)e->tsz);
  if (cpp * ncg > nslab) return m;  // too few slabs for one tile per chunk and case group
  while (cpp * ((nz + L - 1) / L) * ncg > nslab) L *= 2;
  m.mcpp = (int)cpp;
  m.mL = (int)L;
  m.mnzc = (int)((nz + L - 1) / L);
  m.mnt = m.mcpp * m.mnzc;
  m.mncg = (int)ncg;
  return m;
}

// Dynamic LDS of a plane-march launch: the need rounded up to a class, so
// that RVEs of different widths share launch shapes and graphs.  Per CU
// (160 KB, static LDS of at most 3 KB besides): four workgroups up to 40 KB,
// three up to 53 KB.  The first class (38 KB with the static part, a 30^3
// RVE needs 34.4 KB) still fits four but leaves 10 KB of the CU's LDS to the
// other stream group's kernel B workgroups: 0.7% more tangents/s on the
// 128 x 30^3 batch than rounding it up to 40 KB.
static int march_lds_class(int need)
{
  const int cls[4] = {38912 - 3072, 40960 - 3072, 54272 - 3072, kMarchMaxLds};
  for (int c : cls)
    if (need <= c) return c;
  return kMarchMaxLds;
}

// Dynamic LDS of the plane march for an RVE (0: the RVE goes slab by slab).
static int march_lds_need(const mm_engine *e, const HostSlot &h)
{
  const MarchTiling m = march_tiling(e, h.req.rve, h.nslab);
  return m.mcpp > 0 ? march_lds_bytes(e->SN, h.req.rve.nex + 1, h.req.rve.ney + 1, e->K, (int)e->tsz, e->dict_mode)
                     : 0;
}

static void fill_state(const mm_engine *e, const HostSlot &h, RveDev *st, int warm)
{
  memset(st, 0, sizeof(*st));
  const mm_rve_spec &s = h.req.rve;
  st->nex = s.nex;
  st->ney = s.ney;
  st->nez = s.nez;
  st->nx = s.nex + 1;
  st->ny = s.ney + 1;
  st->nz = s.nez + 1;
  st->nrow = h.nrow;
  st->nelm = (int)h.nelm;
  st->nslab = (int)h.nslab;
  st->first_slab = (int)h.first_slab;
  st->row_base = h.first_slab * e->rows_per_slab();
  st->elem_base = h.elem_base;
  st->hx = s.lx / s.nex;
  st->hy = s.ly / s.ney;
  st->hz = s.lz / s.nez;
  st->vol = s.lx * s.ly * s.lz;
  st->nfib = s.nfib;
  for (int i = 0; i < s.nfib; i++) {
    st->fx[i] = s.fib_x[i];
    st->fy[i] = s.fib_y[i];
    st->fr[i] = s.fib_r[i];
  }
  for (int i = 0; i < kNmat; i++) {
    st->young[i] = s.young[i];
    st->poisson[i] = s.poisson[i];
  }
  st->tangent_mode = h.req.tangent_mode;
  st->nload = h.req.nload;
  st->warm = warm;
  st->affine = (e->cfg.initial_guess == MM_GUESS_AFFINE) ? 1 : 0;
  const MarchTiling mt = march_tiling(e, s, h.nslab);
  st->mcpp = mt.mcpp;
  st->mL = mt.mL;
  st->mnzc = mt.mnzc;
  st->mnt = mt.mnt;
  st->mncg = mt.mncg;
  st->big = h.nslab >= e->cfg.big_slabs ? 1 : 0;
  for (int k = 0; k < h.req.nload; k++)
    for (int i = 0; i < MM_NVOI; i++) st->load[k][i] = h.req.strain[k][i];
  if (warm)
    warm_coefficients(h.prev_load, std::min(h.prev_nload, e->K), h.req.strain, h.req.nload, st->wcoef,
                      st->dload);
  for (int c = 0; c < 3; c++) st->st[c].phase = PH_RES;
}

// Admits queued requests that fit; uploads their state and assembles them.
static int admit(mm_engine *e)
{
  std::vector<int> fresh, reused;
  {
    std::lock_guard<std::mutex> lk(e->qmu);
    while (!e->queue.empty()) {
      Queued &qd = e->queue.front();
      const mm_request &req = qd.req;
      // a resident RVE cached under the same key and geometry is reused
      if (req.warm_key) {
        auto it = e->warm_map.find(req.warm_key);
        if (it != e->warm_map.end()) {
          HostSlot &h = e->slots[it->second];
          if (h.state == SL_ACTIVE) break;  // same key still in flight: keep FIFO order
          if (h.state == SL_CACHED && same_geometry(h.req.rve, req.rve)) {
            memcpy(h.prev_load, h.req.strain, sizeof(h.prev_load));
            h.prev_nload = h.req.nload;
            h.req = req;
            h.state = SL_ACTIVE;
            h.t_submit = qd.t_submit;
            h.t_admit = e->now();
            reused.push_back(it->second);
            e->queue.pop_front();
            continue;
          }
          release_slot_resources(e, it->second);  // stale geometry under this key
        }
      }
      int nrow;
      long long nslab, nelm;
      footprint(e, req.rve, &nrow, &nslab, &nelm);
      long long fs = -1, fe = -1;
      int slot = -1;
      for (;;) {
        if (fs < 0) fs = e->slabs.alloc(nslab);
        if (fe < 0) fe = e->elems.alloc(nelm);
        if (slot < 0 && !e->free_slots.empty()) {
          slot = e->free_slots.back();
          e->free_slots.pop_back();
        }
        if (fs >= 0 && fe >= 0 && slot >= 0) break;
        if (!evict_one(e)) break;
      }
      if (fs < 0 || fe < 0 || slot < 0) {
        if (fs >= 0) e->slabs.release(fs, nslab);
        if (fe >= 0) e->elems.release(fe, nelm);
        if (slot >= 0) e->free_slots.push_back(slot);
        break;  // wait for resident RVEs to finish
      }
      HostSlot &h = e->slots[slot];
      h.state = SL_ACTIVE;
      h.req = req;
      h.first_slab = fs;
      h.nslab = nslab;
      h.elem_base = fe;
      h.nelm = nelm;
      h.nrow = nrow;
      h.t_submit = qd.t_submit;
      h.t_admit = e->now();
      if (req.warm_key) e->warm_map[req.warm_key] = slot;
      fresh.push_back(slot);
      e->queue.pop_front();
    }
  }
  const size_t n = fresh.size() + reused.size();
  if (n == 0) return 0;
  int rc = 0;
  // staging: states, slot ids, Ke work, phase-map blocks, slab lists
  size_t nke = fresh.size() * kNmat, nph = 0, nsl = 0, nsl_re = 0;
  for (int s : fresh) {
    nph += (e->slots[s].nelm + 255) / 256;
    nsl += e->slots[s].nslab;
  }
  for (int s : reused) nsl_re += e->slots[s].nslab;
  if ((rc = ensure_pinned(&e->st_states, &e->cap_states, n))) return rc;
  if ((rc = ensure_pinned(&e->st_slots, &e->cap_slots, n))) return rc;
  if ((rc = ensure_pinned(&e->st_ke, &e->cap_ke, 2 * nke + 2))) return rc;
  if ((rc = ensure_pinned(&e->st_phase, &e->cap_phase, nph + 1))) return rc;
  if ((rc = ensure_pinned(&e->st_slabs, &e->cap_slabs, nsl + nsl_re + 1))) return rc;
  if ((rc = ensure_device(&e->dv_states, &e->dcap_states, n))) return rc;
  if ((rc = ensure_device(&e->dv_slots, &e->dcap_slots, n))) return rc;
  if ((rc = ensure_device(&e->dv_ke, &e->dcap_ke, 2 * nke + 2))) return rc;
  if ((rc = ensure_device(&e->dv_phase, &e->dcap_phase, nph + 1))) return rc;
  if ((rc = ensure_device(&e->dv_slabs, &e->dcap_slabs, nsl + nsl_re + 1))) return rc;
  size_t i = 0, ik = 0, ip = 0, is = 0;
  for (int s : fresh) {
    const HostSlot &h = e->slots[s];
    fill_state(e, h, &e->st_states[i], 0);
    e->st_slots[i++] = s;
    for (int m = 0; m < kNmat; m++) {
      e->st_ke[2 * ik] = s;
      e->st_ke[2 * ik + 1] = m;
      ik++;
    }
    for (long long e0 = 0; e0 < h.nelm; e0 += 256) e->st_phase[ip++] = make_int2(s, (int)e0);
    for (long long b = 0; b < h.nslab; b++) e->st_slabs[is++] = (int)(h.first_slab + b);
  }
  for (int s : reused) {
    const HostSlot &h = e->slots[s];
    fill_state(e, h, &e->st_states[i], h.req.warm_start ? 1 : 0);
    e->st_slots[i++] = s;
    for (long long b = 0; b < h.nslab; b++) e->st_slabs[is++] = (int)(h.first_slab + b);
  }
  for (int s : fresh) e->out_host[s].phase = 0;
  for (int s : reused) e->out_host[s].phase = 0;
  MM_HIP(hipMemcpyAsync(e->dv_states, e->st_states, n * sizeof(RveDev), hipMemcpyHostToDevice, e->stream));
  MM_HIP(hipMemcpyAsync(e->dv_slots, e->st_slots, n * sizeof(int), hipMemcpyHostToDevice, e->stream));
  if (nke)
    MM_HIP(hipMemcpyAsync(e->dv_ke, e->st_ke, 2 * nke * sizeof(int), hipMemcpyHostToDevice, e->stream));
  if (nph)
    MM_HIP(hipMemcpyAsync(e->dv_phase, e->st_phase, nph * sizeof(int2), hipMemcpyHostToDevice, e->stream));
  if (nsl + nsl_re)
    MM_HIP(hipMemcpyAsync(e->dv_slabs, e->st_slabs, (nsl + nsl_re) * sizeof(int), hipMemcpyHostToDevice,
                          e->stream));
  EngPtrs P = e->ptrs();
  if ((rc = eng_launch_admit((int)n, e->dv_slots, e->dv_states, e->rves, e->slab_rve, e->slab_node0, e->SN,
                             e->stream)))
    return rc;
  if ((rc = eng_launch_ke((int)nke, e->dv_ke, e->rves, e->ke_tab, e->c_tab, e->bw_tab, e->stream)))
    return rc;
  if ((rc = eng_launch_phase((int)nph, e->dv_phase, e->rves, e->elem_mat, e->stream))) return rc;
  if ((rc = eng_launch_dict(true, (int)fresh.size(), e->dv_slots, e->ke_tab, e->dict64, e->stream))) return rc;
  if (e->dict_mode && !e->fp64 &&
      (rc = eng_launch_dict(false, (int)fresh.size(), e->dv_slots, e->ke_tab, e->dict, e->stream)))
    return rc;
  if ((rc = eng_launch_assemble(e->fp64, e->K, P, (int)nsl, e->dv_slabs, e->dict_mode ? ASM_DICT : ASM_FULL,
                                e->stream)))
    return rc;
  if ((rc = eng_launch_assemble(e->fp64, e->K, P, (int)nsl_re, e->dv_slabs + nsl, ASM_X_ONLY, e->stream)))
    return rc;
  e->n_active += (int)n;
  e->act_dirty = true;
  return 0;
}

// Slabs launched by the stand-alone timing kernels: every allocated slab,
// rounded up to a multiple of 64 (whole waves of workgroups per XCD), never
// beyond the arena, and always a multiple of 8 so the XCD-aware slab order is
// a permutation of the grid.
static int solver_grid(const mm_engine *e, long long hw)
{
  long long g = ((hw + 63) / 64) * 64;
  if (g > e->nslab_cap) g = e->nslab_cap;
  return (int)g;
}

// Rebuilds the lists of slabs being solved (ascending, so each XCD walks
// contiguous slabs) and queues their upload on the engine stream.  With two
// solver streams the RVEs, in slab order, are cut where group 0 first holds
// at least half of the slabs.  The pinned staging buffer is free: the
// previous step ended with a stream sync.
static int update_active(mm_engine *e)
{
  if (!e->act_dirty) return 0;
  std::vector<std::pair<long long, int>> runs;  // (first slab, slot)
  long long total = 0;
  for (int s = 0; s < e->rve_cap; s++)
    if (e->slots[s].state == SL_ACTIVE) {
      runs.push_back({e->slots[s].first_slab, s});
      total += e->slots[s].nslab;
    }
  std::sort(runs.begin(), runs.end());
  const size_t stride = (size_t)e->nslab_cap + 1;
  const int ngroup = e->stream2 ? 2 : 1;
  const bool march = e->cfg.march_planes > 0;
  long long n[2] = {0, 0}, nt[2] = {0, 0};
  int lds[2] = {0, 0}, slab_mode[2] = {0, 0}, big[2] = {0, 0};
  int g = 0;
  for (auto &r : runs) {
    const HostSlot &h = e->slots[r.second];
    if (g == 0 && ngroup == 2 && n[0] > 0 && 2 * n[0] >= total) g = 1;
    for (long long b = 0; b < h.nslab; b++) e->st_act[g * stride + 1 + n[g]++] = (int)(r.first + b);
    if (h.nslab >= e->cfg.big_slabs) big[g] = 1;
    const MarchTiling m = march_tiling(e, h.req.rve, h.nslab);
    if (march && m.mcpp > 0) {
      for (long long t = 0; t < m.tiles(); t++) e->st_act[(2 + g) * stride + 1 + nt[g]++] = (int)(r.first + t);
      lds[g] = std::max(lds[g], march_lds_need(e, h));
    } else {
      slab_mode[g] = 1;
    }
  }
  for (int q = 0; q < 2; q++) {
    e->st_act[q * stride] = (int)n[q];
    e->st_act[(2 + q) * stride] = (int)nt[q];
    if (q < ngroup) {
      MM_HIP(hipMemcpyAsync(e->act_list + q * stride, e->st_act + q * stride, (size_t)(n[q] + 1) * 4,
                            hipMemcpyHostToDevice, e->stream));
      if (march)
        MM_HIP(hipMemcpyAsync(e->act_list + (2 + q) * stride, e->st_act + (2 + q) * stride,
                              (size_t)(nt[q] + 1) * 4, hipMemcpyHostToDevice, e->stream));
    }
    e->act_n[q] = n[q];
    e->tile_n[q] = nt[q];
    e->a_lds[q] = (march && nt[q] > 0) ? march_lds_class(lds[q]) : 0;
    e->a_slab[q] = slab_mode[q];
    e->s_big[q] = big[q];
  }
  e->act_dirty = false;
  return 0;
}

static int launch_grid(long long n)
{
  if (n <= 0) return 0;
  int grid = 64;
  while (grid < n) grid *= 2;
  return grid;
}

// Queues steps_per_launch solver steps (A, B, R each) for every RVE being
// solved: group 0 on the engine stream; group 1, if any, on stream2, forked
// after group 0's first kernel A and joined before the engine stream's other
// work.  The second chain fills the first one's launch gaps and kernel tails
// (A's register use leaves no room for B's waves beside it, so the two do not
// otherwise overlap).
static int queue_steps(mm_engine *e, int grid0, int grid1)
{
  EngPtrs P0 = e->ptrs(), P1 = P0;
  P1.act_list = e->act_list + ((size_t)e->nslab_cap + 1);
  P1.act_tile = e->act_list + 3 * ((size_t)e->nslab_cap + 1);
  P0.a_grid = launch_grid(e->tile_n[0]);
  P1.a_grid = launch_grid(e->tile_n[1]);
  P0.a_lds = e->a_lds[0];
  P1.a_lds = e->a_lds[1];
  P0.a_slab = e->a_slab[0];
  P1.a_slab = e->a_slab[1];
  P0.s_big = e->s_big[0];
  P1.s_big = e->s_big[1];
  const SolveParams sp = e->params();
  const int nk = 3 * e->cfg.steps_per_launch;
  const bool f64 = e->fp64, dm = e->dict_mode;
  int rc = 0;
  if (e->cfg.profile_step_kernel && grid1 == 0) {
    const int ns = e->cfg.steps_per_launch;
    while ((int)e->prof_ev.size() < 2 * ns) {
      hipEvent_t ev;
      MM_HIP(hipEventCreate(&ev));
      e->prof_ev.push_back(ev);
    }
    for (int i = 0; i < ns; i++) {
      MM_HIP(hipEventRecord(e->prof_ev[2 * i], e->stream));
      if ((rc = eng_launch_steps(f64, e->K, dm, P0, sp, grid0, 0, 1, e->stream))) return rc;
      MM_HIP(hipEventRecord(e->prof_ev[2 * i + 1], e->stream));
      if ((rc = eng_launch_steps(f64, e->K, dm, P0, sp, grid0, 1, 2, e->stream))) return rc;
    }
    e->prof_pending = ns;
    return 0;
  }
  if (grid1 == 0) return eng_launch_steps(f64, e->K, dm, P0, sp, grid0, 0, nk, e->stream);
  int first = 0;
  if (grid0 > 0) {
    if ((rc = eng_launch_steps(f64, e->K, dm, P0, sp, grid0, 0, 1, e->stream))) return rc;
    first = 1;
  }
  MM_HIP(hipEventRecord(e->ev_fork, e->stream));
  MM_HIP(hipStreamWaitEvent(e->stream2, e->ev_fork, 0));
  if (grid0 > 0 && (rc = eng_launch_steps(f64, e->K, dm, P0, sp, grid0, first, nk - first, e->stream))) return rc;
  if ((rc = eng_launch_steps(f64, e->K, dm, P1, sp, grid1, 0, nk, e->stream2))) return rc;
  MM_HIP(hipEventRecord(e->ev_join, e->stream2));
  MM_HIP(hipStreamWaitEvent(e->stream, e->ev_join, 0));
  return 0;
}

static int run_steps(mm_engine *e)
{
  int rc = update_active(e);
  if (rc) return rc;
  const int grid0 = launch_grid(e->act_n[0]), grid1 = launch_grid(e->act_n[1]);
  if (grid0 == 0 && grid1 == 0) return 0;
  if (!e->cfg.use_graph) return queue_steps(e, grid0, grid1);
  const std::array<int, 10> key = {grid0, grid1, launch_grid(e->tile_n[0]), launch_grid(e->tile_n[1]),
                                   e->a_lds[0], e->a_lds[1], e->a_slab[0], e->a_slab[1], e->s_big[0],
                                   e->s_big[1]};
  auto it = e->graphs.find(key);
  if (it == e->graphs.end()) {
    MM_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    rc = queue_steps(e, grid0, grid1);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    MM_HIP(ce);
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      return MM_ERR_HIP;
    }
    it = e->graphs.emplace(key, std::make_pair(g, x)).first;
  }
  MM_HIP(hipGraphLaunch(it->second.second, e->stream));
  return 0;
}

static void retire(mm_engine *e)
{
  for (int s = 0; s < e->rve_cap; s++) {
    HostSlot &h = e->slots[s];
    if (h.state != SL_ACTIVE) continue;
    volatile RveOut *o = e->out_host + s;
    if (o->phase != PH_DONE) continue;
    mm_result res;
    memset(&res, 0, sizeof(res));
    res.id = h.req.id;
    res.status = o->status;
    res.pcg_its = o->its_total;
    res.sweeps = o->sweeps;
    res.residual_rel = o->residual_rel;
    for (int k = 0; k < MM_MAX_RHS; k++)
      for (int i = 0; i < MM_NVOI; i++) {
        res.stress[k][i] = o->stress[k][i];
        res.strain[k][i] = o->strain[k][i];
      }
    for (int i = 0; i < 36; i++) res.c_tangent[i] = o->c[i];
    res.t_submit = h.t_submit;
    res.t_admit = h.t_admit;
    res.t_done = e->now();
    o->phase = 0;
    e->n_active--;
    e->act_dirty = true;
    e->completed++;
    if (h.req.warm_key) {
      h.state = SL_CACHED;
      h.lru = ++e->lru_clock;
    } else {
      release_slot_resources(e, s);
    }
    std::lock_guard<std::mutex> lk(e->qmu);
    e->results.push_back(res);
  }
}

extern "C" int mm_engine_step(mm_engine *e)
{
  if (!e) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  MM_HIP(hipSetDevice(e->cfg.device));
  int rc = admit(e);
  if (rc) return e->last_error = rc;
  if (e->n_active > 0) {
    rc = run_steps(e);
    if (rc) return e->last_error = rc;
  }
  MM_HIP(hipStreamSynchronize(e->stream));
  for (int i = 0; i < e->prof_pending; i++) {
    float ms = 0.0f;
    MM_HIP(hipEventElapsedTime(&ms, e->prof_ev[2 * i], e->prof_ev[2 * i + 1]));
    e->prof_ms.push_back(ms);
  }
  e->prof_pending = 0;
  e->nsteps++;
  retire(e);
  return 0;
}

extern "C" int mm_engine_poll(mm_engine *e, mm_result *out, int max, int *n)
{
  if (!e || !n || (max > 0 && !out)) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->qmu);
  int c = 0;
  while (c < max && !e->results.empty()) {
    out[c++] = e->results.front();
    e->results.pop_front();
  }
  *n = c;
  return 0;
}

static bool has_work(mm_engine *e)
{
  std::lock_guard<std::mutex> lk(e->qmu);
  return !e->queue.empty() || e->n_active > 0;
}

extern "C" int mm_engine_drain(mm_engine *e)
{
  if (!e) return MM_ERR_ARG;
  long long stall = 0;
  size_t last_q = (size_t)-1;
  // every RVE finishes within max_sweeps * (max_inner_its + 3) PCG steps;
  // a drain that makes no progress for longer than that reports an error
  const long long limit =
      (long long)e->cfg.max_sweeps * (e->cfg.max_inner_its + 4) / e->cfg.steps_per_launch + 64;
  long long idle = 0, done0 = e->completed;
  while (has_work(e)) {
    int rc = mm_engine_step(e);
    if (rc) return rc;
    if (e->completed != done0) {
      done0 = e->completed;
      idle = 0;
    } else if (++idle > limit) {
      fprintf(stderr, "madmax_amd: drain made no progress in %lld steps\n", idle);
      return MM_ERR_STATE;
    }
    size_t qn;
    int na;
    {
      std::lock_guard<std::mutex> lk(e->qmu);
      qn = e->queue.size();
      na = e->n_active;
    }
    // nothing resident and the head of the queue still does not fit
    if (na == 0 && qn > 0 && qn == last_q) {
      if (++stall > 2) return MM_ERR_FULL;
    } else {
      stall = 0;
    }
    last_q = qn;
  }
  return 0;
}

static void worker_main(mm_engine *e)
{
  (void)hipSetDevice(e->cfg.device);
  while (e->running.load()) {
    if (!has_work(e)) {
      std::unique_lock<std::mutex> lk(e->qmu);
      e->qcv.wait_for(lk, std::chrono::microseconds(200),
                      [&] { return !e->queue.empty() || !e->running.load(); });
      continue;
    }
    if (mm_engine_step(e)) {
      e->running.store(false);
      break;
    }
  }
}

extern "C" int mm_engine_start(mm_engine *e)
{
  if (!e) return MM_ERR_ARG;
  if (e->running.load()) return 0;
  e->running.store(true);
  e->worker = std::thread(worker_main, e);
  return 0;
}

extern "C" int mm_engine_stop(mm_engine *e)
{
  if (!e) return MM_ERR_ARG;
  if (e->running.exchange(false)) e->qcv.notify_all();
  if (e->worker.joinable()) e->worker.join();
  return e->last_error;
}

extern "C" double mm_engine_now(const mm_engine *e) { return e ? e->now() : 0.0; }

extern "C" int mm_engine_info_get(mm_engine *e, mm_engine_info *info)
{
  if (!e || !info) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->qmu);
  info->resident = e->n_active;
  info->queued = (int)e->queue.size();
  info->completed = e->completed;
  info->steps = e->nsteps;
  info->arena_bytes = e->arena_bytes;
  long long free_slabs = 0;
  for (auto &kv : e->slabs.free_) free_slabs += kv.second;
  info->arena_used = (e->slabs.cap > 0)
                         ? (size_t)((double)e->arena_bytes * (double)(e->slabs.cap - free_slabs) /
                                    (double)e->slabs.cap)
                         : 0;
  info->slab_nodes = e->SN;
  info->spmv_ms = e->spmv_ms;
  return 0;
}

extern "C" int mm_engine_step_kernel_times(mm_engine *e, double *ms, int max, int *n)
{
  if (!e || !n || (max > 0 && !ms)) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  const int c = std::min<int>(max, (int)e->prof_ms.size());
  for (int i = 0; i < c; i++) ms[i] = e->prof_ms[i];
  *n = c;
  e->prof_ms.clear();
  return 0;
}

static int find_cached(mm_engine *e, uint64_t key)
{
  auto it = e->warm_map.find(key);
  if (it == e->warm_map.end()) return -1;
  if (e->slots[it->second].state != SL_CACHED) return -1;
  return it->second;
}

extern "C" int mm_engine_rve_sizes(mm_engine *e, uint64_t key, int *nrow, int *nelm)
{
  if (!e) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  const int s = find_cached(e, key);
  if (s < 0) return MM_ERR_STATE;
  if (nrow) *nrow = e->slots[s].nrow;
  if (nelm) *nelm = (int)e->slots[s].nelm;
  return 0;
}

// Expands dictionary storage into the reference's row-major ELL: a boundary
// node's rows are the identity over its class's slots, an interior node's
// rows are its dictionary entry at the structured neighbour columns.
static int download_ell_dict(mm_engine *e, int s, int *cols, double *vals)
{
  const HostSlot &h = e->slots[s];
  const size_t ent = (size_t)kDictPat * kDictEntry;
  std::vector<unsigned char> dict(ent * e->tsz);
  std::vector<uint8_t> pat((size_t)h.nslab * e->SN);
  MM_HIP(hipMemcpyAsync(dict.data(), (char *)e->dict + (size_t)s * ent * e->tsz, ent * e->tsz,
                        hipMemcpyDeviceToHost, e->stream));
  MM_HIP(hipMemcpyAsync(pat.data(), e->npat + (size_t)h.first_slab * e->SN, pat.size(),
                        hipMemcpyDeviceToHost, e->stream));
  MM_HIP(hipStreamSynchronize(e->stream));
  SlotTables T;
  build_slot_tables(&T);
  const mm_rve_spec &rs = h.req.rve;
  const int nx = rs.nex + 1, ny = rs.ney + 1, nz = rs.nez + 1, nxy = nx * ny;
  auto cls_of = [](int i, int n) { return i == 0 ? 0 : (i == n - 1 ? 2 : 1); };
  for (int n = 0; n < h.nrow / 3; n++) {
    const int ix = n % nx, iy = (n / nx) % ny, iz = n / nxy;
    const int cls = cls_of(ix, nx) + 3 * cls_of(iy, ny) + 9 * cls_of(iz, nz);
    const int ns = T.nslot[cls];
    for (int d = 0; d < 3; d++) {
      const size_t row = 3 * (size_t)n + d;
      for (int j = 0; j < kW; j++) {
        const int t = j / 3, dd = j % 3;
        int c = -1;
        double v = 0.0;
        if (t < ns) {
          const int key = T.off[cls][t];
          const int m = n + (key % 3 - 1) + ((key / 3) % 3 - 1) * nx + (key / 9 - 1) * nxy;
          c = 3 * m + dd;
          if (cls != 13) {
            v = (m == n && dd == d) ? 1.0 : 0.0;
          } else {
            const size_t src = (size_t)pat[n] * kDictEntry + (size_t)key * kDictBlk + dd * 4 + d;
            v = e->fp64 ? ((const double *)dict.data())[src] : (double)((const float *)dict.data())[src];
          }
        }
        cols[row * kW + j] = c;
        vals[row * kW + j] = v;
      }
    }
  }
  return 0;
}

extern "C" int mm_engine_download_ell(mm_engine *e, uint64_t key, int *cols, double *vals)
{
  if (!e || !cols || !vals) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  const int s = find_cached(e, key);
  if (s < 0) return MM_ERR_STATE;
  const HostSlot &h = e->slots[s];
  if (e->dict_mode) return download_ell_dict(e, s, cols, vals);
  const size_t RS = (size_t)e->rows_per_slab();
  const size_t nv = (size_t)h.nslab * RS * kW, nc = (size_t)h.nslab * e->SN * 27;
  std::vector<int> hc(nc);
  std::vector<unsigned char> hv(nv * e->tsz);
  MM_HIP(hipMemcpyAsync(hc.data(), e->bcols + (size_t)h.first_slab * e->SN * 27, nc * 4,
                        hipMemcpyDeviceToHost, e->stream));
  MM_HIP(hipMemcpyAsync(hv.data(), (char *)e->vals + (size_t)h.first_slab * RS * kW * e->tsz,
                        nv * e->tsz, hipMemcpyDeviceToHost, e->stream));
  MM_HIP(hipStreamSynchronize(e->stream));
  // block key of a node holds the reference's slot t = key_slot[class][key]:
  // column 3 * bcols[n][key] + d' sits at slot 3t + d' of row 3n + d
  SlotTables T;
  build_slot_tables(&T);
  const mm_rve_spec &rs = h.req.rve;
  const int nx = rs.nex + 1, ny = rs.ney + 1, nz = rs.nez + 1;
  auto cls_of = [](int i, int n) { return i == 0 ? 0 : (i == n - 1 ? 2 : 1); };
  for (size_t i = 0; i < (size_t)h.nrow * kW; i++) {
    cols[i] = -1;
    vals[i] = 0.0;
  }
  for (int n = 0; n < h.nrow / 3; n++) {
    const int ix = n % nx, iy = (n / nx) % ny, iz = n / (nx * ny);
    const int cls = cls_of(ix, nx) + 3 * cls_of(iy, ny) + 9 * cls_of(iz, nz);
    const size_t sl = (size_t)(n / e->SN), lane = n % e->SN;
    for (int key = 0; key < 27; key++) {
      const int t = T.key_slot[cls][key];
      if (t < 0) continue;
      const int m = hc[sl * e->SN * 27 + (size_t)key * e->SN + lane];
      for (int d = 0; d < 3; d++)
        for (int dd = 0; dd < 3; dd++) {
          const size_t dst = (size_t)(3 * n + d) * kW + 3 * t + dd;
          const size_t src = sl * RS * kW + (size_t)(3 * key + dd) * RS + 3 * lane + d;
          cols[dst] = (m < 0) ? -1 : 3 * m + dd;
          vals[dst] = e->fp64 ? ((const double *)hv.data())[src] : (double)((const float *)hv.data())[src];
        }
    }
  }
  return 0;
}

extern "C" int mm_engine_download_phase(mm_engine *e, uint64_t key, int *elem_type)
{
  if (!e || !elem_type) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  const int s = find_cached(e, key);
  if (s < 0) return MM_ERR_STATE;
  const HostSlot &h = e->slots[s];
  std::vector<uint8_t> m(h.nelm);
  MM_HIP(hipMemcpyAsync(m.data(), e->elem_mat + h.elem_base, h.nelm, hipMemcpyDeviceToHost, e->stream));
  MM_HIP(hipStreamSynchronize(e->stream));
  for (long long i = 0; i < h.nelm; i++) elem_type[i] = m[i];
  return 0;
}

extern "C" int mm_engine_download_x(mm_engine *e, uint64_t key, int load, double *x)
{
  if (!e || !x) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  if (load < 0 || load >= e->K) return MM_ERR_ARG;
  const int s = find_cached(e, key);
  if (s < 0) return MM_ERR_STATE;
  const HostSlot &h = e->slots[s];
  const size_t RS = (size_t)e->rows_per_slab();
  std::vector<double> hx((size_t)h.nslab * RS * e->K);
  MM_HIP(hipMemcpyAsync(hx.data(), e->x64 + (size_t)h.first_slab * RS * e->K, hx.size() * sizeof(double),
                        hipMemcpyDeviceToHost, e->stream));
  MM_HIP(hipStreamSynchronize(e->stream));
  // [slab][K][3*SN] -> row-major for one load case
  for (int i = 0; i < h.nrow; i++) {
    const size_t sl = (size_t)i / RS, lr = (size_t)i % RS;
    x[i] = hx[(sl * e->K + load) * RS + lr];
  }
  return 0;
}

// Times `reps` launches of a stand-alone solver kernel over the resident RVEs
// with HIP events on the engine stream.  what 0: the SpMV w = A z; what 1:
// kernel A's INNER work (SpMV and the p, q update).  Algorithmic bytes of one
// launch: z read once per row and load case, w written (what 0) or p and q
// read and written (what 1); full storage adds every row's 81 values and
// every node's 27 block columns, dictionary storage one pattern byte per node
// and each resident RVE's dictionary once.
static int time_kernel(mm_engine *e, int what, int reps, double *avg_ms, double *bytes)
{
  if (!e || reps < 1) return MM_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->step_mu);
  long long hw = e->slabs.high_water();
  if (hw <= 0 || e->n_active > 0) return MM_ERR_STATE;  // it overwrites solver vectors
  const int grid = solver_grid(e, hw);
  EngPtrs P = e->ptrs();
  double rows = 0.0, nrve = 0.0;
  int lds = 0, slab_mode = 0;
  for (int s = 0; s < e->rve_cap; s++)
    if (e->slots[s].state != SL_FREE) {
      rows += e->slots[s].nrow;
      nrve += 1.0;
      const int need = march_lds_need(e, e->slots[s]);
      lds = std::max(lds, need);
      if (need == 0) slab_mode = 1;
    }
  if (e->cfg.march_planes > 0 && lds > 0) {
    P.a_lds = march_lds_class(lds);
    P.a_slab = slab_mode;
  }
  hipEvent_t a, b;
  MM_HIP(hipEventCreate(&a));
  MM_HIP(hipEventCreate(&b));
  int rc = eng_launch_time(e->fp64, e->K, e->dict_mode, P, grid, what, e->stream);  // warm-up
  if (rc) return rc;
  MM_HIP(hipEventRecord(a, e->stream));
  for (int i = 0; i < reps; i++) {
    rc = eng_launch_time(e->fp64, e->K, e->dict_mode, P, grid, what, e->stream);
    if (rc) return rc;
  }
  MM_HIP(hipEventRecord(b, e->stream));
  MM_HIP(hipEventSynchronize(b));
  float ms = 0.0f;
  MM_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (what == 0) e->spmv_ms = ms / reps;
  if (avg_ms) *avg_ms = ms / reps;
  if (bytes) {
    const double vec = rows * (what == 0 ? 2.0 : 5.0) * e->K * e->tsz;
    *bytes = e->dict_mode ? vec + rows / 3.0 + nrve * kDictPat * kDictEntry * (double)e->tsz
                          : vec + rows * (kW * (double)e->tsz + 27.0 * 4.0 / 3.0);
  }
  return 0;
}

extern "C" int mm_engine_time_spmv(mm_engine *e, int reps, double *avg_ms, double *bytes)
{
  return time_kernel(e, 0, reps, avg_ms, bytes);
}

extern "C" int mm_engine_time_step(mm_engine *e, int reps, double *avg_ms, double *bytes)
{
  return time_kernel(e, 1, reps, avg_ms, bytes);
}

extern "C" int mm_hbm_copy_probe(int device, size_t bytes, int reps, double *avg_ms)
{
  if (bytes == 0 || reps < 1 || !avg_ms) return MM_ERR_ARG;
  MM_HIP(hipSetDevice(device));
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float ms = 0.0f;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto done;
  }
  if (hipMemsetAsync(a, 1, bytes, s) != hipSuccess ||
      hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipEventRecord(e0, s) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto done;
  }
  for (int i = 0; i < reps; i++)
    if (hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      rc = MM_ERR_HIP;
      goto done;
    }
  if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto done;
  }
  *avg_ms = ms / reps;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// ---------------------------------------------------------------------------
// Hand-written streaming probes: the practical HBM ceiling of this device for
// the two access patterns of the solver kernels.  Grid-stride loops over
// 16-byte non-temporal accesses, 8 workgroups of 256 lanes per CU.
namespace {

typedef float stream_f4 __attribute__((ext_vector_type(4)));

// dst = src (reads and writes `n` 16-byte words)
__global__ __launch_bounds__(256) void k_stream_copy(const stream_f4 *__restrict__ src, stream_f4 *__restrict__ dst,
                                                     long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// reads `n` 16-byte words; one float per workgroup is written so that the
// loads are not dead
__global__ __launch_bounds__(256) void k_stream_read(const stream_f4 *__restrict__ src, float *__restrict__ out,
                                                     long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  float acc = 0.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const stream_f4 v = __builtin_nontemporal_load(src + i);
    acc += (v.x + v.y) + (v.z + v.w);
  }
  __shared__ float red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace

extern "C" int mm_hbm_stream_probe(int device, size_t bytes, int reps, int mode, double *avg_ms)
{
  if (bytes < 16 || reps < 1 || !avg_ms || (mode != 0 && mode != 1)) return MM_ERR_ARG;
  MM_HIP(hipSetDevice(device));
  int ncu = 0;
  MM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  const long long n = (long long)(bytes / 16);
  const int grid = 8 * (ncu > 0 ? ncu : 256);
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float ms = 0.0f;
  const size_t bbytes = (mode == 0) ? (size_t)n * 16 : (size_t)grid * sizeof(float);
  if (hipMalloc(&a, (size_t)n * 16) != hipSuccess || hipMalloc(&b, bbytes) != hipSuccess ||
      hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto done;
  }
  if (hipMemsetAsync(a, 0, (size_t)n * 16, s) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto done;
  }
  for (int i = 0; i <= reps; i++) {  // launch 0 is the warm-up
    if (i == 1 && hipEventRecord(e0, s) != hipSuccess) {
      rc = MM_ERR_HIP;
      goto done;
    }
    if (mode == 0) k_stream_copy<<<grid, 256, 0, s>>>((const stream_f4 *)a, (stream_f4 *)b, n);
    else k_stream_read<<<grid, 256, 0, s>>>((const stream_f4 *)a, (float *)b, n);
  }
  if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
      hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
    rc = MM_ERR_HIP;
    goto done;
  }
  *avg_ms = ms / reps;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}
