/*
 * f2m.h -- C ABI of the MI355X fractional 2-matching (F2M) solver.
 *
 * The reference (aav-antonov/F2M) has no library or FFI: its stable surface
 * is the CLI `F2M.solver.exe <tsplib> <K>` (RUN.cpp:12-31) and the internal
 * C++ seam  graph(file,K) -> graph_c(graph*) -> graph::RUN()
 * (F2M.h:175,181,320).  Each entry point below names the part of that seam
 * it replaces.  A maintainer binding this from another language (ctypes,
 * cgo, JNI) needs only these declarations: plain pointers, sizes and PODs.
 *
 * Conventions
 *  - Status codes: 0 ok, < 0 error (the reference calls exit(); this ABI
 *    never does).  f2m_strerror() names a code, f2m_last_error() gives the
 *    message of the last failing call on a context.
 *  - Array arguments marked [dev|host] may be device memory or host memory;
 *    host memory is staged through the context.  [dev] must be device memory.
 *  - A context owns all device workspace.  Calls on one context are ordered on
 *    its stream and are not thread-safe; use one context per host thread.
 *    Device arrays returned through f2m_graph / f2m_batch stay valid until the
 *    next build call on the same context or f2m_ctx_destroy().
 *  - Node ids are the input order after exact-duplicate removal.  Each node's
 *    slots are sorted by (cost, neighbour id); an undirected edge {u,w}
 *    appears once in u's slots and once in w's.
 */
#ifndef F2M_H
#define F2M_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t f2m_status;
#define F2M_OK 0
#define F2M_ERR_ARG (-1)        /* bad argument                                   */
#define F2M_ERR_HIP (-2)        /* HIP runtime error (reference: cudaCheckErrors,
                                   graph_cuda_define.hpp:12-22, exit(1))           */
#define F2M_ERR_FORMAT (-3)     /* not a TSPLIB coordinate file (FXY.hpp:221-226)  */
#define F2M_ERR_IO (-4)         /* file cannot be opened or written               */
#define F2M_ERR_DEGENERATE (-5) /* no positive edge cost / fewer than 2 points     */
#define F2M_ERR_LIMIT (-6)      /* size beyond a documented limit                  */

/* Edge-cost modes.  RAW_FP64 is the reference's (FXY.hpp:382: unrounded double
 * distance).  NINT = floor(dist + 0.5).  RAW_FP32 = dist rounded to float. */
#define F2M_COST_RAW_FP64 0
#define F2M_COST_NINT 1
#define F2M_COST_RAW_FP32 2

/* Solve schedule; f2m_default_opts() fills the reference's hard-coded values. */
typedef struct f2m_opts {
  double err_grad;        /* 1e-7  convergence of max|g|   graph_cuda.cu:56         */
  double err_comb_mult;   /* 1e2   eps = err_grad*mult     graph_cuda.cu:57         */
  int32_t glob_k;         /* 5     warm-start list length  graph_cuda.cu:370        */
  int32_t glob_rounds;    /* 200                           graph_cuda.cu:375        */
  int32_t glob_sweeps;    /* 50                            graph_cuda.cu:377        */
  double glob_exit;       /* 0.01                          graph_cuda.cu:394        */
  int32_t max_outer;      /* 10000                         graph_cuda.cu:178        */
  int32_t factor0;        /* 100                           graph_cuda.cu:177        */
  int32_t factor_step;    /* 300                           graph_cuda.cu:186        */
  int32_t factor_max;     /* 5000                          graph_cuda.cu:186        */
  int32_t check_every;    /* 50                            graph_cuda.cu:323        */
  int32_t early_exit;     /* 0 reference block_EF; 1 break on the current max|g|  */
  int32_t repair;         /* 1 run the dual-perturbation loop graph_cuda.cu:69-126 */
  double shift0;          /* 0.5                           graph_cuda.cu:72         */
  double shift_max;       /* 32                            graph_cuda.cu:124        */
  int32_t fix_rounds;     /* 100                           graph_cuda.cu:214        */
  int32_t fix_factor;     /* 500                           graph_cuda.cu:216        */
  int32_t frozen_sweeps;  /* 20                            graph_cuda_comb.hpp:437  */
  int32_t stop_phase;     /* 0 full; 1 after warm start; 2 after dual descent      */
  /* Exact primal recovery (new; replaces the least squares + perturbation of
   * comb_thread.cpp:272-304 / graph_cuda.cu:69-126 where they fail): warm-
   * started cost-scaling push-relabel on the bipartite double cover, ending
   * with a certified duality gap (below 1/2 for integer costs: exact).
   * x is then in {0, 1/2, 1} with degree exactly 2. */
  int32_t exact;          /* 1     0 never (the reference's output); 1 when the
                                   reference recovery is not exact (status 1, or
                                   an irregular tight component, whose x would
                                   come from float least squares); 2 always.
                                   With 1 or 2 the least squares of irregular
                                   components and the dual repairs are not run
                                   (n_irregular_b = n_repairs = 0): the exact
                                   stage replaces them                           */
  int32_t exact_alpha;    /* 4     eps divisor between phases                     */
  double exact_eps0;      /* 0.02  first-phase eps, units of d_min                */
  double exact_tol;       /* 1e-10 certified relative duality gap                 */
  int64_t exact_max_rounds; /* 1e6 push/relabel rounds over all phases            */
  int32_t exact_max_phases; /* 40                                                 */
  int32_t exact_reserved;
} f2m_opts;

typedef struct f2m_stats {
  double objective;       /* sum_e x_e cost_e in FP64 (graph_tour.hpp:165-184)     */
  double dual;            /* d_min * (-sum F + sum_e min(0, d_e))                  */
  double max_grad;        /* final max |g| (normalised units)                      */
  double d_min;           /* smallest positive edge cost                           */
  int64_t sweeps;         /* Jacobi sweeps executed                                */
  int32_t glob_rounds;
  int32_t outer_rounds;
  int32_t n_combs;        /* tight components with >= 2 nodes                      */
  int32_t n_irregular_a;  /* components failing 2A+B=4                             */
  int32_t n_irregular_b;  /* ... whose least-squares solve did not converge        */
  int32_t n_repairs;      /* dual-perturbation attempts                            */
  int32_t status;         /* 0 "Optimal Found: Yes", 1 "ERROR: Optimal Solution
                             not Found" (graph_tour.hpp:218-241)                   */
  int32_t max_rank;       /* MAX edge rank (graph_tour.hpp:169-180)                */
  int32_t exact_ran;      /* 1 if the exact recovery ran (x, F, comb_root are then
                             its output: comb_root = components of x = 1/2 edges) */
  int32_t exact_phases;   /* eps phases                                            */
  int64_t exact_rounds;   /* push/relabel rounds with an active node               */
  double exact_gap;       /* certified primal - dual gap, raw cost units           */
  int32_t exact_certified;/* 1 if the gap is within tolerance                      */
  int32_t exact_reserved;
} f2m_stats;

/* Device CSR graph (replaces graph + graph_c::init/init_cuda,
 * graph_tour.hpp:2-110 and graph_cuda_init.hpp:3-193). */
typedef struct f2m_graph {
  int32_t n;              /* nodes                                                 */
  int32_t k;              /* neighbours per node used to build it (0: user CSR)    */
  int32_t cost_mode;
  int32_t reserved;
  int64_t nslots;         /* 2E                                                    */
  double d_min;           /* smallest positive cost (graph_tour.hpp:94)            */
  const int32_t* rowptr;  /* [dev] n+1                                             */
  const int32_t* nbr;     /* [dev] nslots                                          */
  const double* cost;     /* [dev] nslots, raw cost                                */
} f2m_graph;

/* A batch of independent instances packed back to back (new; no reference
 * counterpart).  Instance i owns nodes [node_off[i], node_off[i+1]) and slots
 * [rowptr[node_off[i]], rowptr[node_off[i+1]]); node ids in nbr are local to
 * the instance. */
typedef struct f2m_batch {
  int32_t n_inst;
  int32_t k;
  int32_t cost_mode;
  int32_t reserved;
  int64_t n_total;
  int64_t nslots;
  const int64_t* node_off; /* [dev] n_inst+1                                       */
  const int64_t* rowptr;   /* [dev] n_total+1, global slot offsets                 */
  const int32_t* nbr;      /* [dev] nslots, instance-local node ids                */
  const double* cost;      /* [dev] nslots                                         */
} f2m_batch;

typedef struct f2m_ctx f2m_ctx;
typedef struct f2m_points f2m_points;

void f2m_default_opts(f2m_opts* o);
const char* f2m_strerror(f2m_status s);
const char* f2m_last_error(const f2m_ctx* ctx);
/* Library build identifier, e.g. "f2m_hip gfx950 <date>". */
const char* f2m_version(void);

/* Context on HIP device `device`; `stream` is a hipStream_t (NULL: the context
 * creates its own non-blocking stream at default priority). */
f2m_status f2m_ctx_create(int32_t device, void* stream, f2m_ctx** out);
void f2m_ctx_destroy(f2m_ctx* ctx);
f2m_status f2m_ctx_synchronize(f2m_ctx* ctx);

/* ---- instrumentation (new; the reference prints gettimeofday deltas,
 * graph_cuda_define.hpp:29-34) -------------------------------------------------
 * With sample_every > 0, the sweep launches of f2m_solve of the dominant
 * kind are timed with HIP events on the context stream: on spatial tiles
 * (F2M_FUSE sweeps per launch, default 10, on the block_EF lists) every run
 * of two or more back-to-back launches of that width, by one event pair
 * around the run (the gaps between its launches included; the sweeps'
 * bits do not depend on it); with F2M_FUSE=1 every sample_every-th single
 * sweep, by events the runtime records at the kernel's own start and end.
 * f2m_get_profile reports the timed launches and their time, the sweeps
 * and algorithmic bytes of those launches, and the algorithmic bytes of
 * all sweeps:
 * 8 B per active slot (nbr int32 + a 4-byte cost) + 20 B per node (list
 * offset, F in, F out).  The solver stores D as a double (12 B per slot
 * moved); the count is the algorithm's, not the layout's. */
typedef struct f2m_profile {
  int64_t sweep_launches;   /* sweep kernel launches of the last solve         */
  int64_t sweep_samples;    /* launches timed with events                      */
  double sweep_ms_sampled;  /* summed duration of the timed launches (ms)      */
  double sweep_alg_bytes;   /* algorithmic bytes of all sweeps of the solve    */
  int64_t active_slots;     /* sum over sweeps of the active slots processed   */
  double solve_ms;          /* wall time of the last f2m_solve (host clock)    */
  int64_t sweeps;           /* sweeps executed (a launch may run several)      */
  int64_t sampled_sweeps;   /* sweeps run by the timed launches                */
  double sampled_alg_bytes; /* algorithmic bytes of those sweeps               */
} f2m_profile;
f2m_status f2m_ctx_set_profile(f2m_ctx* ctx, int32_t sample_every);
f2m_status f2m_get_profile(f2m_ctx* ctx, f2m_profile* out);

/* ---- TSPLIB input (replaces get_size + FXY::get_KXY, FXY.hpp:186-309) ------
 * Reads NODE_COORD_SECTION up to EOF, rounds coordinates to 1e-3, drops exact
 * duplicate points (kept label -> dropped labels).  Host-side. */
f2m_status f2m_tsplib_read(const char* path, f2m_points** out);
int64_t f2m_points_count(const f2m_points* p);
int64_t f2m_points_duplicates(const f2m_points* p);
/* Copies labels/x/y (any may be NULL) into caller host arrays of count(). */
void f2m_points_copy(const f2m_points* p, int64_t* labels, double* x, double* y);
/* Writes `<path>` in the reference's _duplicates format (FXY.hpp:168-183). */
f2m_status f2m_points_write_duplicates(const f2m_points* p, const char* path);
void f2m_points_free(f2m_points* p);

/* ---- init: kNN edge construction (replaces FXY::FXY kNN + graph::graph
 * selection, FXY.hpp:23-144 and graph_tour.hpp:28-98) ------------------------
 * x, y [dev|host] n doubles.  Builds the union over v of the K nearest
 * neighbours of v under the key (cost, id), as a CSR on the device. */
f2m_status f2m_knn_build(f2m_ctx* ctx, int32_t n, const double* x, const double* y, int32_t k,
                         int32_t cost_mode, f2m_graph* out);
/* Spatial numbering (new; the reference numbers nodes in unordered_map order,
 * graph_tour.hpp:97).  perm[i] = input index of the i-th point along the Morton
 * (Z-order) curve: coordinates quantised to 16 bits per axis over the bounding
 * box, q = min(65535, floor(((v - min) / (max - min)) * 65535)) (0 on a zero
 * extent), x in the even key bits, ties in input order.  Building the graph on
 * the points permuted this way (x'[i] = x[perm[i]]) puts graph neighbours close
 * in memory and makes f2m_solve faster (n = 100k Gaussian mixture: 63.7 ms
 * against 79.3 ms in input order); node i of that graph is input point
 * perm[i].  x, y [dev|host] n doubles; perm [dev|host] n int32. */
f2m_status f2m_spatial_order(f2m_ctx* ctx, int32_t n, const double* x, const double* y, int32_t* perm);
/* The same per instance of a packed batch (node_off [host] n_inst+1): perm
 * [dev|host] node_off[n_inst] global input indices, each instance's range
 * permuted within itself. */
f2m_status f2m_spatial_order_batched(f2m_ctx* ctx, int32_t n_inst, const int64_t* node_off, const double* x,
                                     const double* y, int64_t* perm);
/* Directed kNN only: out [dev|host] n*k int32, row v = v's k nearest sorted. */
f2m_status f2m_knn_directed(f2m_ctx* ctx, int32_t n, const double* x, const double* y,
                            int32_t k, int32_t cost_mode, int32_t* out);
/* Wrap a caller CSR (copied into the context); slots must be sorted as above. */
f2m_status f2m_graph_from_csr(f2m_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* nbr,
                              const double* cost, int64_t nslots, f2m_graph* out);

/* Copy a graph's CSR to caller arrays [dev|host] (n+1, nslots, nslots); any
 * may be NULL.  Replaces reading graph::V / graph::E (F2M.h:120-188). */
f2m_status f2m_graph_copy(f2m_ctx* ctx, const f2m_graph* g, int32_t* rowptr, int32_t* nbr, double* cost);

/* ---- solve (replaces graph::RUN, graph_cuda.cu:47-155) ----------------------
 * Dual descent with the reference schedule, tight components, primal
 * recovery, objective.  Results stay on the device; copy with f2m_result_copy. */
f2m_status f2m_solve(f2m_ctx* ctx, const f2m_graph* g, const f2m_opts* o, f2m_stats* st);
/* Warm start (new; the reference always starts from F = 0 in block_GLOB,
 * graph_cuda.cu:363-413): the same schedule, block_GLOB included, from
 * potentials F_init [dev|host] n (normalised, as f2m_result_copy returns them,
 * e.g. of the same points with a smaller k, or moved slightly).  Skipping
 * block_GLOB as well is not offered: from a perturbed optimum the Jacobi
 * descent then failed to converge on some instances. */
f2m_status f2m_solve_warm(f2m_ctx* ctx, const f2m_graph* g, const f2m_opts* o, const double* F_init,
                          f2m_stats* st);
/* kNN restriction certificate (new; the reference reports the MAX edge rank,
 * graph_tour.hpp:169-180, as a hint).  After a solve of g (built by
 * f2m_knn_build with k), counts the nodes v with r_v < -d_min F_v, r_v = cost of
 * v's k-th nearest neighbour.  An edge left out of the graph costs at least
 * max(r_u, r_w), so with 0 such nodes every left-out edge has a nonnegative
 * reduced cost: the dual stays feasible over all pairs and the optimum on the
 * graph is the LP optimum over the complete graph.  (Sufficient, not
 * necessary: a graph can be optimal with nodes counted.) */
f2m_status f2m_knn_certify(f2m_ctx* ctx, const f2m_graph* g, int64_t* violations);
/* F [n] potentials, x [nslots] per canonical slot, comb_root [n] smallest node
 * id of the node's tight component or -1.  Any pointer may be NULL; [dev|host]. */
f2m_status f2m_result_copy(f2m_ctx* ctx, double* F, double* x, int32_t* comb_root);

/* ---- kernel-level entry points of the dual descent (graph_cuda_grad.hpp,
 * graph_cuda_partition.hpp, graph_cuda_maxmin.hpp), used by the parity tests --
 * One or more Jacobi sweeps over the first len[v] slots of each node, slots in
 * the given order: g = -(t2+t3)/2, F <- F + g.  D [dev|host] normalised costs
 * per slot; len [dev|host] n; F_in [dev|host] n; F_out, g_out [dev|host] n.
 * max_abs_g receives max_v |g_v| of the last sweep (may be NULL). */
f2m_status f2m_sweep(f2m_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* nbr,
                     const double* D, const int32_t* len, const double* F_in, double* F_out,
                     double* g_out, int32_t nsweeps, double* max_abs_g);

/* ---- allocation-free sweeps (new; no counterpart in the reference) -----------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops and harnesses that forbid hipMalloc.  The same sweeps
 * and results as f2m_sweep, bit for bit, with every array a [dev] array of the
 * caller (host arrays are rejected, F2M_ERR_ARG) and F_in != F_out; g_out may
 * be NULL; max_abs_g [dev|host] may be NULL (a device pointer keeps the call
 * free of host synchronisation).  No hipMalloc, hipFree or other allocation:
 * the intermediate potentials and the max reduction use a fixed device array
 * of the library (8 MiB + 4 KiB per device, held until the process ends).
 * n above F2M_NOALLOC_MAX_N fails with F2M_ERR_LIMIT.  All calls on one device
 * share that array: calls that overlap on a device (two streams, two threads)
 * race, so serialise them. */
#define F2M_NOALLOC_MAX_N 1048576
f2m_status f2m_sweep_noalloc(f2m_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* nbr,
                             const double* D, const int32_t* len, const double* F_in, double* F_out,
                             double* g_out, int32_t nsweeps, double* max_abs_g);

/* ---- batched form (new) ------------------------------------------------------
 * node_off [host] n_inst+1 offsets into x, y [dev|host]; one kNN graph per
 * instance, all with the same k and cost mode. */
f2m_status f2m_knn_build_batched(f2m_ctx* ctx, int32_t n_inst, const int64_t* node_off,
                                 const double* x, const double* y, int32_t k, int32_t cost_mode,
                                 f2m_batch* out);
/* Solves every instance with the plain schedule, all in one persistent launch
 * per phase.  st [host] n_inst stats (may be NULL). */
f2m_status f2m_solve_batched(f2m_ctx* ctx, const f2m_batch* b, const f2m_opts* o, f2m_stats* st);
/* Warm start of every instance (new): F_init [dev|host] n_total potentials
 * (normalised, as f2m_batch_result_copy returns them, e.g. of the previous
 * forward of a training loop), as f2m_solve_warm. */
f2m_status f2m_solve_batched_warm(f2m_ctx* ctx, const f2m_batch* b, const f2m_opts* o, const double* F_init,
                                  f2m_stats* st);
/* Autotuner (new): times the schedules (lanes per node 1/2/4/8 x workgroup
 * 256/512/1024 threads) on the first `sample` instances (per-instance device
 * clocks, weighted by workgroup size), keeps the fastest per size class
 * (<=500, 1000, 2000, 4000, 8000, 16000, larger) for later f2m_solve_batched
 * calls on this context.  With cache_path, an existing file ("max_n lanes
 * threads placement" per line; placement -1 = default, else 0-3 as
 * F2M_BATCH_LDS) is loaded instead of tuning, else the chosen table is
 * written there (placement -1).  The schedules differ in speed only: results
 * are bit-identical. */
f2m_status f2m_batch_autotune(f2m_ctx* ctx, const f2m_batch* b, const f2m_opts* o, int32_t sample,
                              const char* cache_path);
/* Copy a batch's arrays to caller arrays [dev|host]: node_off (n_inst+1),
 * rowptr (n_total+1), nbr and cost (nslots); any may be NULL. */
f2m_status f2m_batch_copy(f2m_ctx* ctx, const f2m_batch* b, int64_t* node_off, int64_t* rowptr, int32_t* nbr,
                          double* cost);
/* x [dev|host] nslots per canonical slot, F [dev|host] n_total. */
f2m_status f2m_batch_result_copy(f2m_ctx* ctx, double* F, double* x, double* objective);
/* comb_root [dev|host] n_total: per node of the last f2m_solve_batched, the
 * smallest instance-local node id of its tight component, or -1 (as
 * f2m_result_copy's comb_root; with the exact stage: components of the
 * x = 1/2 edges).  Replaces reading graph_c's comb_id / comb_V arrays
 * (graph_cuda_comb.hpp:201-273) per instance. */
f2m_status f2m_batch_comb_root_copy(f2m_ctx* ctx, int32_t* comb_root);

/* ---- differentiable bound (new) -------------------------------------------
 * dV/dp for V = sum_e x_e |p_u - p_w| (envelope theorem: x* is fixed).
 * x_slot [dev] per canonical slot (f2m_solve / f2m_solve_batched output),
 * px, py [dev] float per node, gx, gy [dev] float per node, scaled by
 * grad_scale[instance] (NULL: 1).  Batch form: node_off/rowptr from f2m_batch. */
f2m_status f2m_bound_backward(f2m_ctx* ctx, const f2m_batch* b, const double* x_slot,
                              const float* px, const float* py, const float* grad_scale,
                              float* gx, float* gy);

#ifdef __cplusplus
}
#endif
#endif
