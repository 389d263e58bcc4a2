/*
 * f2m_oracle.h -- CPU restatement of the reference F2M hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the f2m_amd product links, loads or
 * calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function restates a piece of the reference (aav-antonov/F2M) and
 * cites the file:line it follows.  Conventions shared with the product
 * (node ids, slot order, tie rule, class boundaries) are documented at each
 * function.  Parity pins: see tests/test_oracle.py (golden LP optima from
 * HiGHS, brute-force kNN, closed-form sweeps, strong duality).
 */
#ifndef F2M_ORACLE_H
#define F2M_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge-cost modes.  RAW_FP64: unrounded double Euclidean distance
 * (FXY.hpp:382, the reference's only mode).  NINT: floor(dist + 0.5).
 * RAW_FP32: dist rounded to float. */
enum { ORC_COST_RAW_FP64 = 0, ORC_COST_NINT = 1, ORC_COST_RAW_FP32 = 2 };

/* Schedule constants; defaults equal the reference's hard-coded values. */
typedef struct orc_opts {
  double err_grad;        /* 1e-7   graph_cuda.cu:56 (error_B)               */
  double err_comb_mult;   /* 1e2    graph_cuda.cu:57 (error_B_comb)          */
  int32_t glob_k;         /* 5      graph_cuda.cu:370                        */
  int32_t glob_rounds;    /* 200    graph_cuda.cu:375                        */
  int32_t glob_sweeps;    /* 50     graph_cuda.cu:377                        */
  double glob_exit;       /* 0.01   graph_cuda.cu:394                        */
  int32_t max_outer;      /* 10000  graph_cuda.cu:178                        */
  int32_t factor0;        /* 100    graph_cuda.cu:177                        */
  int32_t factor_step;    /* 300    graph_cuda.cu:186                        */
  int32_t factor_max;     /* 5000   graph_cuda.cu:186                        */
  int32_t check_every;    /* 50     graph_cuda.cu:323                        */
  int32_t early_exit;     /* 0: reference block_EF (break tests the entry
                             max|g|, graph_cuda.cu:329); 1: test the
                             current max|g| at each check                   */
  int32_t repair;         /* 1: run the dual-perturbation loop
                             (graph_cuda.cu:69-126)                          */
  double shift0;          /* 0.5    graph_cuda.cu:72                         */
  double shift_max;       /* 32     graph_cuda.cu:124                        */
  int32_t fix_rounds;     /* 100    graph_cuda.cu:214                        */
  int32_t fix_factor;     /* 500    graph_cuda.cu:216                        */
  int32_t frozen_sweeps;  /* 20     graph_cuda_comb.hpp:437                  */
  int32_t stop_phase;     /* 0 full; 1 after block_GLOB; 2 after dual descent */
  /* exact primal recovery (no reference counterpart, f2m_exact_oracle.cpp) */
  int32_t exact;          /* 0 never; 1 when the reference recovery is not
                             exact (status 1 or an irregular tight
                             component); 2 always                            */
  int32_t exact_alpha;    /* 4      eps divisor between phases               */
  double exact_eps0;      /* 0.02   first-phase eps, units of d_min          */
  double exact_tol;       /* 1e-10  certified relative duality gap           */
  int64_t exact_max_rounds; /* 1e6  push/relabel rounds over all phases      */
  int32_t exact_max_phases; /* 40                                           */
  int32_t exact_reserved;
} orc_opts;

typedef struct orc_stats {
  double objective;       /* sum x_e * cost_e, raw units (graph_tour.hpp:165-184) */
  double dual;            /* -sum F + sum min(0,d), times d_min                  */
  double max_grad;        /* final max |g| over active lists (normalised)        */
  double d_min;           /* smallest positive edge cost                         */
  int64_t sweeps;         /* Jacobi sweeps executed                              */
  int32_t glob_rounds;
  int32_t outer_rounds;
  int32_t n_combs;
  int32_t n_irregular_a;  /* combs failing 2A+B=4 (comb_thread.cpp:180-195)      */
  int32_t n_irregular_b;  /* of those, least squares did not converge            */
  int32_t n_repairs;      /* perturbation attempts                               */
  int32_t status;         /* 0 optimum found, 1 ERROR (graph_tour.hpp:236)       */
  int32_t max_rank;       /* MAX edge rank (graph_tour.hpp:169-180)              */
  int32_t exact_ran;      /* 1 if the exact recovery ran                         */
  int32_t exact_phases;   /* eps phases                                          */
  int64_t exact_rounds;   /* push/relabel rounds with an active node             */
  double exact_gap;       /* certified P - D, raw units                          */
  int32_t exact_certified;/* gap within tolerance                                */
  int32_t exact_reserved;
} orc_stats;

void orc_default_opts(orc_opts* o);

/* --- init: kNN candidate selection (FXY.hpp:441-534, graph_tour.hpp:28-98) --- */
double orc_cost(double xu, double yu, double xw, double yw, int mode);
/* Directed kNN: out[v*K + j] = j-th nearest of v under key (cost, id). */
int orc_knn_directed(int n, const double* x, const double* y, int K, int mode,
                     int brute, int nthreads, int32_t* out);
/* Symmetrised graph, CSR with each node's slots sorted by (cost, nbr). */
void* orc_graph_build(int n, const double* x, const double* y, int K, int mode,
                      int brute, int nthreads);
int64_t orc_graph_nslots(void* h);
void orc_graph_export(void* h, int32_t* rowptr, int32_t* nbr, double* cost);
void orc_graph_free(void* h);

/* --- grad: one damped Jacobi sweep over the first len[v] slots -------------
 * (graph_cuda_grad.hpp:1-12, 54-61, 65-84) */
void orc_sweep(int n, const int32_t* rowptr, const int32_t* nbr, const double* D,
               const int32_t* len, const double* F_in, double* F_out, double* g_out);

/* --- partition (graph_cuda_partition.hpp:3-54) on one list of d values.
 * Permutes idx[] in place exactly as the reference permutes its pointers;
 * returns the pivot. raise!=0 selects qpart_ptr_EF (threshold +1 until >= 3). */
int orc_qpart(double* d, int32_t* idx, int size, double thr, int raise);

/* --- full solve (graph::RUN, graph_cuda.cu:47-155) ---------------------------
 * x_out is per canonical slot (input order), comb_root[v] = smallest node id
 * of v's tight component, or -1 if v has no tight edge. */
int orc_solve(int n, const int32_t* rowptr, const int32_t* nbr, const double* cost,
              const orc_opts* o, int nthreads, double* F_out, double* x_out,
              int32_t* comb_root, orc_stats* st);

/* --- warm start (new): the same schedule (block_GLOB included,
 * graph_cuda.cu:363-413) from potentials F_init (normalised, as F_out)
 * instead of zero. */
int orc_solve_warm(int n, const int32_t* rowptr, const int32_t* nbr, const double* cost,
                   const orc_opts* o, int nthreads, const double* F_init, double* F_out, double* x_out,
                   int32_t* comb_root, orc_stats* st);

/* --- kNN restriction certificate (new; the reference reports the MAX edge
 * rank, graph_tour.hpp:169-180, as a hint).  Counts the nodes v with
 * r_v < -d_min F_v, r_v = cost of v's K-th nearest neighbour (slot K-1 of its
 * (cost, id)-sorted row; nodes of degree < K are skipped).  An edge (u, w) left
 * out of the symmetrised kNN graph has cost >= max(r_u, r_w); with y = -d_min
 * F / 2 its reduced cost c - y_u - y_w is then >= 0 whenever no node is
 * counted, so the dual stays feasible on the complete graph and the optimum on
 * the graph is the optimum over all pairs. */
int64_t orc_knn_certify(int n, const int32_t* rowptr, const double* cost, int K, double dmin, const double* F);

/* --- least squares of the irregular combs (Ax_b.cpp:19-158) ----------------
 * rows: CSR over equations (row_ptr[m+1], col[nnz]); all coefficients 1.
 * x_out[nv]; returns iterations used (go_max) or 0 if not converged. */
int orc_axb(int m, const int32_t* row_ptr, const int32_t* col, const int32_t* b,
            int nv, double* x_out, double* max_err);

/* --- exact primal recovery (f2m_exact_oracle.cpp; no reference counterpart) --
 * Warm-started cost-scaling push-relabel on the bipartite double cover from
 * potentials F_in (normalised, as orc_solve's F).  Writes x per canonical
 * slot, the certifying potentials F_out, half_root[v] = smallest id of v's
 * component of x = 1/2 edges (-1 if none) and the exact_* stats.  Returns 0
 * if certified, 1 if not, -1 on an asymmetric graph, -2 on a node of degree
 * below 2. */
int orc_exact(int n, const int32_t* rowptr, const int32_t* nbr, const double* cost, double dmin,
              const double* F_in, const orc_opts* o, double* F_out, double* x_out, int32_t* half_root,
              orc_stats* st);

/* --- backward: dV/dp for V = sum x_e |p_u - p_w| (envelope theorem) -------- */
void orc_bound_backward(int n, const int32_t* rowptr, const int32_t* nbr,
                        const double* x_slot, const double* px, const double* py,
                        double* gx, double* gy);

#ifdef __cplusplus
}
#endif
#endif
