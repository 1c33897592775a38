/*
 * micro_api.h -- MI355X-native drop-in for the reference's micro (RVE)
 * homogenization entry points (micro/inc/micro.h:150-157) plus the batched /
 * serving engine behind them.
 *
 * Conventions (3D hex8, straight extensions of the reference's 2D code):
 *   node  n = ix + iy*nx + iz*nx*ny                (src/mesh.c:869-876)
 *   elem  e = ex + ey*nex + ez*nex*ney             (src/mesh.c:878-888)
 *   elem dof index k*3 + d -> node_k*3 + d          (src/mesh.c:890-901)
 *   Voigt [xx, yy, zz, xy, yz, xz], engineering shears (micro/src/init.c:40-50)
 *   c_tangent row-major, c[j*6 + i] = d sigma_j / d eps_i
 *                                                   (micro/src/homogenize.c:80)
 *   Boundary condition USTRAIN: every node on the six faces gets u = E.X
 *                                                   (micro/src/homogenize.c:227-242)
 * Return codes follow the reference's int convention: 0 ok, non-zero error.
 */
#ifndef MADMAX_MICRO_API_H_
#define MADMAX_MICRO_API_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_NVOI 6
#define MM_MAX_FIBERS 8
#define MM_MAX_RHS 8
#define HOMOGENIZE_DELTA_STRAIN 0.005 /* micro/inc/micro.h:29 */

/* status / error codes */
#define MM_OK 0
#define MM_ERR_ARG 1
#define MM_ERR_ALLOC 2
#define MM_ERR_FULL 3
#define MM_NOT_CONVERGED 5
#define MM_ERR_HIP 6
#define MM_ERR_STATE 7
#define MM_REJECTED 8 /* larger than the whole HBM budget */

/* boundary conditions of the RVE (the values of micro/inc/micro.h:26) */
#define MICRO_BC_USTRAIN 1 /* u = E X on the six faces */
#define MICRO_BC_USTRESS 2 /* only sized by the reference (micro/src/alloc.c:18-19): refused */
#define MICRO_BC_PER_LM 3  /* periodic: corners u = E X, u+ - u- = E (X+ - X-) on opposite faces */

/* solver choice for the process-global micro */
#define MICRO_SOLVER_PCG 0       /* FP64 Jacobi-PCG (PETSc cg + jacobi semantics) */
#define MICRO_SOLVER_PCG_MIXED 1 /* FP32 Jacobi-PCG inside FP64 iterative refinement */

/* An RVE: structured nex x ney x nez hex8 box of size lx x ly x lz with
 * z-aligned cylindrical fibers (micro/src/micro_struct.c:98-129: an element is
 * fiber when its centroid's in-plane distance to a fiber centre is <= r). */
typedef struct {
  int nex, ney, nez;
  double lx, ly, lz;
  int nfib;
  double fib_x[MM_MAX_FIBERS], fib_y[MM_MAX_FIBERS], fib_r[MM_MAX_FIBERS];
  double young[2], poisson[2]; /* [0] MATRIX, [1] FIBER (src/material.c:84-135) */
} mm_rve_spec;

/* ---------------------------------------------------------------------------
 * Process-global micro, as micro/src/main.c:57-109 sets it up (argv parsing
 * is replaced by this struct).  micro_init assembles the RVE on the GPU and,
 * since all materials are linear, computes c_tangent_linear with
 * homog_calculate_c_tangent_around_zero like init_variables_2
 * (micro/src/init.c:54-59). */
typedef struct {
  mm_rve_spec rve;
  int solver;          /* MICRO_SOLVER_* */
  double rtol;         /* PCG relative residual target (default 1e-10) */
  int nl_max_its;      /* Newton / refinement sweeps (micro/src/init.c:7, default 2) */
  double nl_min_norm;  /* absolute residual threshold (micro/src/init.c:8, default 1e-4) */
  double rho;          /* density answered to ACTION_MICRO_CALC_RHO (micro/src/init.c:9) */
  int device;
  int linear_materials; /* 1 (default): the tangent is computed once at init and
                           stresses come from it (micro/src/homogenize.c:3-21);
                           0: flags.linear_materials == false, every request
                           runs homog_fe2's Newton loop with the Jacobian
                           re-assembled element by element on the GPU at every
                           iteration (micro/src/homogenize.c:23-49, 175-225) */
  int fe2_bc;           /* MICRO_BC_USTRAIN (default) or MICRO_BC_PER_LM */
} micro_config;

int micro_config_default(micro_config *cfg);
int micro_init(const micro_config *cfg);
int micro_finalize(void);

/* micro/inc/micro.h:154 -- linear materials answer with c_tangent_linear */
int homog_get_strain_stress(double *strain_mac, double *strain_ave, double *stress_ave);
/* micro/inc/micro.h:155 -- always runs the RVE solve (homog_fe2) */
int homog_get_strain_stress_non_linear(double *strain_mac, double *strain_ave,
                                       double *stress_ave);
/* micro/inc/micro.h:157 -- one RVE solve for one macro strain */
int homog_fe2(double *strain_mac, double *strain_ave, double *stress_ave);
/* micro/inc/micro.h:150 -- points at library-owned storage */
int homog_get_c_tangent(double *strain_mac, double **c_tangent);
/* micro/inc/micro.h:152 -- forward differences, 7 solves done as one
 * 7-right-hand-side block on the GPU */
int homog_calculate_c_tangent(double *strain_mac, double *c_tangent);
/* micro/inc/micro.h:153 */
int homog_calculate_c_tangent_around_zero(double *c_tangent);

/* The request / response record of the macro-micro coupling, byte-for-byte
 * the layout of message_t (inc/comm.h:18-27), and the dispatch of one request
 * as the micro's loop does it (micro/src/main.c:116-151).  Returns 0, or 1
 * for an unknown action. */
#define ACTION_NULL 0
#define ACTION_MICRO_CALC_STRESS 1
#define ACTION_MICRO_CALC_C_TANGENT 2
#define ACTION_MICRO_CALC_RHO 3
#define ACTION_MICRO_END 4
typedef struct {
  int action;
  int num_voigt;
  double strain_mac[6];
  double stress_ave[6];
  double c_tangent_ave[36];
  double rho;
} micro_message;
int micro_serve_message(micro_message *msg);

/* Inspection of the process-global micro (no reference counterpart). */
typedef struct {
  int pcg_its;      /* inner iterations of the last solve (max over RHS) */
  int sweeps;       /* Newton / refinement sweeps of the last solve */
  double residual;  /* final FP64 residual norm (max over RHS) */
  double seconds;   /* wall time of the last solve */
  int assemblies;   /* Jacobian assemblies since micro_init (element path) */
  int pad_;
} micro_stats;
int micro_get_stats(micro_stats *st);
int micro_get_sizes(int *nrow, int *nnz, int *nelm);
/* assembled Jacobian in the reference's row-major ELL layout */
int micro_get_ell(int *cols, double *vals);
int micro_get_elem_type(int *elem_type);
/* displacement of the last homog_fe2 call (length nrow; with
 * MICRO_BC_PER_LM the periodic fluctuation on the nex x ney x nez nodes of
 * the torus, u = E X + x) */
int micro_get_displacement(double *x);
/* assembly_res_ell (micro/src/assembly.c:33-82) on the GPU for a given
 * displacement x (length nrow): r = sum_e sum_gp B^T sigma w with the
 * boundary dofs zeroed, and |r|_2.  With MICRO_BC_PER_LM, x is the periodic
 * fluctuation and strain the macro strain (NULL: zero) of the affine part. */
int micro_residual(const double *strain, const double *x, double *r, double *norm);
/* Writes <name>.pvtu and <name>_0.vtu for the last homog_fe2 call like
 * micro_pvtu (micro/src/output.c:43-188): node positions, hex8 cells,
 * displacement and residual per node, part, strain, stress (element averages
 * over the gauss points) and elem_type per cell. */
int micro_write_vtu(const char *name);

/* ---------------------------------------------------------------------------
 * Batched / serving engine: many RVEs of uneven size resident in one HBM
 * arena, each with a block of load cases solved together. */

#define MM_PREC_FP64 0  /* FP64 PCG */
#define MM_PREC_MIXED 1 /* FP32 PCG inside FP64 iterative refinement */

/* How the batched solver reads the ELL matrix of an RVE.  Both give the same
 * SpMV bits.  FULL streams the per-row values and block columns from HBM (the
 * reference's layout, slab-blocked).  DICT exploits the structured two-phase
 * mesh: the 3-row block of an interior node is fixed by the materials of its
 * 8 elements, so each RVE keeps a dictionary of the 256 possible blocks (the
 * same FP64 sums, rounded) and one pattern byte per node; the matrix then
 * lives in L1/L2 instead of HBM. */
#define MM_ELL_FULL 0
#define MM_ELL_DICT 1

/* Interior displacement a cold solve starts from.  ZERO is the reference's
 * start (set_disp_0 sets the boundary only); AFFINE starts every node at the
 * applied affine field u = E X, which is exact for a homogeneous cell and
 * leaves only the residual of the phase interfaces (fewer PCG iterations, the
 * same converged solution; tolerances stay relative to the zero-start
 * right-hand side). */
#define MM_GUESS_ZERO 0
#define MM_GUESS_AFFINE 1

#define MM_TANGENT_NONE 0
#define MM_TANGENT_UNIT 1 /* c[j*6+i] = sigma_i[j] / strain_i[i], loads i = 0..5 */
#define MM_TANGENT_FD 2   /* c[j*6+i] = (sigma_{i+1}[j]-sigma_0[j]) / (eps_{i+1}[i]-eps_0[i]) */

typedef struct {
  int device;
  size_t hbm_budget;     /* bytes for the RVE arena (0 = 80% of free memory) */
  int max_rves;          /* resident RVE slots (0 = 4096) */
  int slab_nodes;        /* nodes per ELL slab = lanes per workgroup: 64, 128 or 256
                            (0 = 128); a slab holds 3 * slab_nodes rows */
  int nrhs;              /* load cases per RVE, 1..MM_MAX_RHS */
  int precision;         /* MM_PREC_* */
  double rtol;           /* FP64 residual target relative to the first residual */
  double abs_tol;        /* absolute FP64 residual threshold (0 = off) */
  double inner_rtol;     /* per-sweep PCG target */
  int max_inner_its;
  int max_sweeps;
  int steps_per_launch;  /* PCG iterations per launch / graph replay */
  int use_graph;         /* capture the steps in a hipGraph */
  int sync_every_step;   /* plain path: host reads the state after every step */
  int ell_storage;       /* MM_ELL_FULL or MM_ELL_DICT (default DICT) */
  int initial_guess;     /* MM_GUESS_ZERO or MM_GUESS_AFFINE (default AFFINE) */
  int solver_streams;    /* 1, or 2 (default): resident RVEs stepped in two
                            groups on two HIP streams */
  int profile_step_kernel; /* 1: no graph, one stream, and every launch of the
                            PCG step kernel (kernel A: SpMV + p/q update)
                            timed with HIP events (mm_engine_step_kernel_times) */
  int march_planes;      /* kernel A: 0 works slab by slab, gathering every
                            neighbour from memory; n > 0 (default 8) marches
                            in-plane chunks of slab_nodes nodes through runs
                            of n z-planes with the neighbour planes staged in
                            LDS (nrhs >= 2 only: two load cases per tile in
                            FP32, one in FP64; RVEs whose x-extent exceeds
                            slab_nodes/2 - 1 elements stay slab by slab).  Same SpMV bits; the p.q
                            partial sums are grouped differently. */
  int big_slabs;         /* RVEs of at least this many slabs (default 1024)
                            have their per-iteration dot products summed once
                            per RVE (kernels R and S) instead of in every
                            workgroup of the kernel that uses them: the same
                            bits, with work linear instead of quadratic in
                            the RVE's slab count */
} mm_engine_config;

typedef struct {
  uint64_t id;
  mm_rve_spec rve;
  int nload;                      /* 1 .. engine nrhs */
  double strain[MM_MAX_RHS][MM_NVOI];
  int tangent_mode;               /* MM_TANGENT_* */
  uint64_t warm_key;              /* 0: nothing kept; else the RVE (mesh, ELL,
                                     displacements) stays resident under this
                                     key and a later request with the same key
                                     and geometry reuses it without assembly */
  int warm_start;                 /* with a reused RVE: 0 starts cold (as the
                                     engine's initial_guess); 1 starts from the
                                     kept displacements: if six kept load cases
                                     span the strain space, load k starts from
                                     their combination that matches its strain
                                     (exact for linear materials), else from
                                     kept load k plus the affine field of the
                                     strain increment */
} mm_request;

typedef struct {
  uint64_t id;
  int status;                     /* MM_OK or MM_NOT_CONVERGED / MM_REJECTED */
  int pcg_its;                    /* total inner iterations (sum over sweeps) */
  int sweeps;
  double residual_rel;            /* final max over RHS of ||r|| / ||r_0|| */
  double stress[MM_MAX_RHS][MM_NVOI];
  double strain[MM_MAX_RHS][MM_NVOI];
  double c_tangent[MM_NVOI * MM_NVOI];
  double t_submit, t_admit, t_done; /* engine clock, seconds */
} mm_result;

typedef struct mm_engine mm_engine;

int mm_engine_config_default(mm_engine_config *cfg);
/* Host-side check of a request for an engine with nrhs load cases (no GPU
 * call; mm_engine_submit runs it first).  Returns MM_ERR_ARG for an invalid
 * RVE, non-finite strains, a load count outside 1..nrhs, and for load cases
 * that do not have the shape the tangent mode divides by: MM_TANGENT_UNIT
 * needs load i = a_i e_i (a_i != 0, every other component exactly 0) for
 * i = 0..5; MM_TANGENT_FD needs load i+1 to differ from load 0 in component
 * i and in no other (micro/src/homogenize.c:59-94). */
int mm_request_validate(const mm_request *req, int nrhs);
int mm_engine_create(const mm_engine_config *cfg, mm_engine **eng);
int mm_engine_destroy(mm_engine *eng);
/* Queues a request (thread-safe); returns MM_REJECTED if it can never fit:
 * its slabs or elements exceed the arena, or one solver vector of it (nrhs x
 * 3 x nodes rounded up to slabs, times 4 or 8 bytes) reaches 2^31 bytes, the
 * reach of the SpMV's 32-bit offsets (about 12.8 M nodes at 7 FP64 loads). */
int mm_engine_submit(mm_engine *eng, const mm_request *req);
/* One scheduling round: admit queued requests that fit, run
 * steps_per_launch PCG steps for every resident RVE, retire finished ones. */
int mm_engine_step(mm_engine *eng);
/* Copies up to max finished results out; *n receives the count. */
int mm_engine_poll(mm_engine *eng, mm_result *out, int max, int *n);
/* Runs mm_engine_step until every queued and resident request is finished. */
int mm_engine_drain(mm_engine *eng);
/* Background worker thread that keeps stepping while work exists. */
int mm_engine_start(mm_engine *eng);
int mm_engine_stop(mm_engine *eng);
/* Engine clock (seconds since creation). */
double mm_engine_now(const mm_engine *eng);

typedef struct {
  int resident;       /* RVEs on the device */
  int queued;
  long long completed;
  long long steps;
  size_t arena_bytes, arena_used;
  int slab_nodes;
  double spmv_ms;     /* average duration of the ELL SpMV kernel, from HIP events */
} mm_engine_info;
int mm_engine_info_get(mm_engine *eng, mm_engine_info *info);
/* Durations (ms) of the step-kernel launches recorded since the last call
 * by an engine created with profile_step_kernel (oldest first, at most max);
 * *n receives the count and the record is cleared. */
int mm_engine_step_kernel_times(mm_engine *eng, double *ms, int max, int *n);

/* Inspection of an RVE kept resident under warm_key (finished requests):
 * the Jacobian in the reference's row-major ELL layout (nrow x 81, values
 * widened to double), the element phases, and the displacement of one load. */
int mm_engine_rve_sizes(mm_engine *eng, uint64_t warm_key, int *nrow, int *nelm);
int mm_engine_download_ell(mm_engine *eng, uint64_t warm_key, int *cols, double *vals);
int mm_engine_download_phase(mm_engine *eng, uint64_t warm_key, int *elem_type);
int mm_engine_download_x(mm_engine *eng, uint64_t warm_key, int load, double *x);

/* Times `reps` launches of the batched ELL SpMV+dot kernel on the resident
 * RVEs with HIP events on the engine stream; writes the average launch
 * duration and the algorithmic bytes of one launch. */
int mm_engine_time_spmv(mm_engine *eng, int reps, double *avg_ms, double *bytes);
/* The same for the PCG step's SpMV kernel as it runs in the solve: SpMV plus
 * the p = z + beta p, q = w + beta q update and the p.q partials (the solve's
 * dominant kernel).  Both overwrite solver vectors: only with no RVE being
 * solved (MM_ERR_STATE otherwise). */
int mm_engine_time_step(mm_engine *eng, int reps, double *avg_ms, double *bytes);

/* Copies `bytes` between two device buffers `reps` times (the runtime's
 * device-to-device copy kernel) and reports the average duration: the
 * practical HBM ceiling, and the known-size reference that calibrates the
 * FETCH_SIZE / WRITE_SIZE counters. */
int mm_hbm_copy_probe(int device, size_t bytes, int reps, double *avg_ms);

/* Hand-written streaming kernels over `bytes` bytes (16-byte non-temporal
 * accesses, grid-stride, 8 workgroups per CU), timed with HIP events over
 * `reps` launches after one warm-up: mode 0 copies (reads and writes `bytes`
 * each), mode 1 only reads.  The HBM rate a kernel of that access pattern
 * can reach on this device. */
int mm_hbm_stream_probe(int device, size_t bytes, int reps, int mode, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
