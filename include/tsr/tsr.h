/*
 * tsr.h - C-ABI of the MI355X-native TwoScaleRich time-stepping library (libtsr.so).
 *
 * Three groups of entry points:
 *   1. Reference signatures, so a caller written for the reference links unchanged.
 *      Each cites the reference declaration it replaces.
 *   2. Problem set-up: the benchmark cases (small, plain, ragged, 64-scenario, large)
 *      and the reference's own input files, producing the batch view (tsr_types.h)
 *      that the device path consumes.
 *   3. A device-resident solver: state stays in HBM between steps.
 *
 * Errors: functions returning int give 0 on success; nonzero codes are listed per
 * function.  tsr_last_error() returns a message for the most recent failure.
 */
#ifndef TSR_H
#define TSR_H

#include "tsr_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ================================ 1. reference signatures ============================== */

/* RHS on the GPU.  Replaces Gfunc_parallel / Gfunc_serial (include/main.h:21-22; selected
 * at main.cpp:198-213).  u and g are host arrays of user_data.N doubles.  The problem is
 * the batch view registered for (user_data.macro_mesh.elements, N) by the set-up calls,
 * or, if none is registered, the reference's semantics (one micro mesh, soil B in every
 * micro element, cells hosted by macro elements when N = M_nodes + M_elements*m_nodes,
 * by macro nodes when N = M_nodes*(1 + m_nodes)).  Returns 0, 1 if g is not finite,
 * 2 on a device or set-up error. */
int Gfunc_hip(double* g, double* u, user_data_struct user_data);

/* One accepted EXPREM step, C twin of the reference's C++ exprem (include/main.h:29-32):
 * t and stats by pointer.  When Gfunc == Gfunc_hip the whole step runs on the device;
 * any other Gfunc is called on the host for every RHS evaluation while the Krylov
 * arithmetic stays on the device.  Returns 0 or -1 (main.cpp:319-322). */
int exprem_c(tsr_gfunc_t Gfunc, int N, double* t, double tend, double* tspan, double* tspan_solns, double* u0,
             user_data_struct user_data, exprem_options_struct options, exprem_stats_struct* stats);

/* phi_1(alpha*H[0:m,0:m]) by the (6,6) Pade approximant with scaling and squaring.
 * src/phipade.cpp:15 (H leading dimension MaxKrylov+1, phiH MaxKrylov).  Returns 0, or 1
 * when the Pade denominator is singular (the reference exits instead, phipade.cpp:162-167). */
int phipade(double* phiH, double* H, double alpha, int m, int MaxKrylov);
/* The solver's phi functions (no counterpart in the reference, whose exprem source is
 * absent).  tsr_phi1_e1: phi_1(alpha*H[0:m,0:m]) e1 (the first column only, into
 * out[0..m-1]; H leading dimension ldh), which is all the step and its Krylov estimate
 * need: with flush_denormals = 0 the same values, bit for bit, as the first column of
 * phipade, at about half its cost for large alpha*H (the doublings form one block of
 * columns of phi instead of all).  tsr_phik_e1: out[(k-1)*m + i] = [phi_k(tau H) e1]_i,
 * k = 1..kmax, from the exponential of the augmented matrix (the error estimate's phi_3).
 * flush_denormals != 0 (the solver's setting): subnormal intermediates are taken as zero
 * (x86 FTZ/DAZ on the calling thread during the call), which changes only values below
 * 2.2e-308 and makes stiff evaluations several times faster.  Return 0, or 1 as phipade. */
int tsr_phi1_e1(double* out, const double* H, double alpha, int m, int ldh, int flush_denormals);
int tsr_phik_e1(double* out, const double* H, int ldh, int m, double tau, int kmax, int flush_denormals);

/* sqrt(mean((PhiError_i / (AbsTol + RelTol |u_i|))^2)) over host arrays.  include/exprem.h:27 */
double weighted_norm(double* PhiError, double* u, double AbsTol, double RelTol, int N);

/* include/exprem.h:29: reports (stderr) a step below MinStep. */
void check_step(double dt, double MinStep);

/* include/exprem.h:30: step size to retry with after the RHS returned flag (halved if
 * flag != 0, unchanged otherwise). */
double check_flag_Gfunc(int flag, double dt);

/* Frees the device copy and the cached solver that Gfunc_hip / Gfunc_* / exprem made for
 * a caller's own user_data (the reference's driver never frees its meshes,
 * main.cpp:357-361, so these live until this call or the process end).  A user_data
 * whose arrays were edited in place or reallocated is noticed by a checksum on the next
 * call and re-uploaded.  No counterpart in the reference.  Returns 0, or 1 when nothing
 * was adopted for this user_data (problems of tsr_problem_* are freed by
 * tsr_problem_free). */
int tsr_release(user_data_struct user_data);

/* Defaults of src/set_exprem_options.cpp:16-26 (C twin of set_exprem_options). */
int set_exprem_options_c(exprem_options_struct* options);

/* Initialise stats before the first step (the reference leaves them unset, main.cpp:257;
 * exprem_c also does this itself when *t == 0). */
void exprem_stats_init(exprem_stats_struct* stats, const exprem_options_struct* options, double* u0);

/* ===================================== 2. set-up ======================================= */

typedef struct tsr_problem tsr_problem;

/* Workloads: 0 plain CPU-runnable (macro 21^2, micro 11^2, t in [0,0.1] d), 1 plain GPU
 * scale (101^2 / 21^2, [0,1] d), 2 ragged cells (101^2; micro 11/16/21/31 per cell, 1-3
 * random inclusions), 3 batch of independent scenarios (64 x (41^2 / 16^2), sampled
 * soils), 4 large (201^2 / 31^2, one random inclusion per cell).  macro_n, micro_r and
 * n_scen override the sizes when > 0 (micro_r is ignored for workload 2).  seed = 0 uses
 * the workload's fixed seed. */
enum { TSR_CFG_PLAIN_SMALL = 0, TSR_CFG_PLAIN = 1, TSR_CFG_RAGGED = 2, TSR_CFG_SCENARIOS = 3, TSR_CFG_LARGE = 4 };
tsr_problem* tsr_problem_config(int cfg, int macro_n, int micro_r, int n_scen, unsigned long long seed);

/* The reference's input files (tests/Richards_Equation/: *.msh, Keff.txt).  host_kind
 * TSR_HOST_ELEMENT reproduces main.cpp:260.  Returns NULL on a read error. */
tsr_problem* tsr_problem_reference(const char* macro_msh, const char* micro_msh, const char* keff_txt,
                                   const soil_struct* soilA, const soil_struct* soilB, double boundary_flux,
                                   const int* macro_bc5, int model_type, double h0, int host_kind);

void                    tsr_problem_free(tsr_problem* p);
const tsr_batch_view*   tsr_problem_view(const tsr_problem* p);
const user_data_struct* tsr_problem_user_data(const tsr_problem* p);   /* N = whole batch */
long long               tsr_problem_N(const tsr_problem* p);
void                    tsr_problem_initial(const tsr_problem* p, double* u0);
double                  tsr_problem_tend(const tsr_problem* p);
void                    tsr_problem_options(const tsr_problem* p, exprem_options_struct* o);
/* Geometry-distinct bytes a G call must read beyond u and g (per-cell inclusion
 * parameters, per-scenario soils and Keff tables); the algorithmic traffic of one G
 * call is 16*N plus this. */
long long               tsr_problem_distinct_bytes(const tsr_problem* p);

/* 1 when a micro template is the uniform r x r right-triangle grid the analytic
 * micro-cell kernel handles (connectivity, Dirichlet perimeter and geometry checked). */
int tsr_template_structured(const micro_mesh_struct* m);

const char* tsr_last_error(void);

/* ============================== 3. device-resident solver ============================== */

typedef struct tsr_solver tsr_solver;

/* Inputs of the learned step controller, per scenario at the start of a step:
 *   0      log10 of the step the classical law proposes now
 *   1..3   log10 of the last three accepted steps (newest first)
 *   4..7   Krylov dimensions of the last four accepted steps (newest first)
 *   8      log10 ||H_m||_inf of the last accepted step
 *   9      log10 of its last Arnoldi residual h_{m+1,m}
 *   10..12 mean, min, max macro head [m]
 *   13..15 mean, min, max effective saturation of the matrix soil at the macro heads
 * Entries without history (first steps) repeat the newest value available, or 0. */
enum { TSR_FEAT_COUNT = 16 };

/* Per-step record (the reference's statistics.txt line, main.cpp:324-329, extended with
 * what the step controller saw). */
typedef struct tsr_step_record {
    double t;            /* time after the step */
    double tau;          /* step taken */
    double tau_proposed; /* step proposed before error control */
    double err;          /* WRMS error estimate of the accepted attempt */
    double beta;         /* ||G(u)||_2 at the step start */
    double hnorm;        /* ||H_m||_inf (Hessenberg of the accepted step) */
    double resid;        /* last Arnoldi residual h_{m+1,m} */
    double head_mean, head_min, head_max, sat_mean, sat_min, sat_max;  /* macro head / saturation stats */
    double tau_first;    /* step of the first attempt (after the Krylov-budget halvings) */
    double err_first;    /* WRMS error estimate of the first attempt (inf if its RHS failed) */
    float  feat[16];     /* the controller's inputs at the step start (order above) */
    double wall_ms;      /* host wall time of the step (to the end of its last sync) */
    int    m;            /* Krylov dimension of phi_1 */
    int    mD;           /* Krylov dimension of the error estimate */
    int    rejects;      /* rejected attempts before acceptance */
    int    fevals;       /* RHS evaluations in this step */
    int    controller;   /* 0 classical, 1 learned proposal, 2 learned output unusable (classical), 3 exploration */
    int    m_cap;        /* Krylov cap used */
    double tau_next;     /* step the classical law proposes next (exprem_stats_struct.CurrentStep) */
} tsr_step_record;

/* Writes (append = 0) or appends one line per record in the reference's statistics.txt
 * format (src/main.cpp:324-329): the step number (first_step, first_step + 1, ...) right-
 * aligned in 5 columns, then t after the step, the step proposed next and the Krylov
 * dimension, "%5d %.4e %.4e %d".  Returns 0, or 1 when the file cannot be opened. */
int tsr_statistics_write(const char* path, const tsr_step_record* recs, int n, int first_step, int append);

typedef struct tsr_run_stats {
    double t;
    int    steps, rejects, fevals, failed;
    double wall_s;
    double gfunc_ms, arnoldi_ms;   /* device time in the RHS and in the Krylov kernels */
    double host_dense_s;           /* host time in the small dense phi routines */
    double sync_wait_s;            /* host time blocked on the device stream */
} tsr_run_stats;

/* Creates device copies of the problem (mesh templates, soils, Keff tables) and the
 * Krylov workspace for opt->MaxKrylov (plus learned-controller cap up to 60).  Returns
 * NULL on failure. */
tsr_solver* tsr_solver_create(const tsr_batch_view* view, const exprem_options_struct* opt);
void        tsr_solver_free(tsr_solver* s);
int         tsr_solver_set_state(tsr_solver* s, const double* u_host, double t);
int         tsr_solver_get_state(tsr_solver* s, double* u_host, double* t);
void        tsr_solver_stats(const tsr_solver* s, exprem_stats_struct* st);
/* Reset statistics and step-size proposal (tau = options.InitStep). */
void        tsr_solver_reset(tsr_solver* s);

/* g = G(u) with u, g host arrays (copies through HBM). */
int tsr_solver_gfunc_host(tsr_solver* s, double* g, const double* u);
/* g_dev = G(u_dev) on device pointers; repeat launches `reps` times (timing). */
int tsr_solver_gfunc_dev(tsr_solver* s, double* g_dev, const double* u_dev, int reps);
/* Device time (ms) of the last tsr_solver_gfunc_dev call per launch, from HIP events on
 * the solver's stream. */
double tsr_solver_last_gfunc_ms(const tsr_solver* s);
/* Launch shape the autotuner chose for the micro-cell kernel on this problem
 * (workgroup size, waves-per-EU register budget; 0 before the first G). */
void tsr_solver_micro_shape(const tsr_solver* s, int* bs, int* wpe);
/* Average device time (ms) of each of G's three parts over `reps` evaluations of
 * G(u_dev), HIP events between the launches: [0] is 0 (the macro elements run in the
 * leading workgroups of the micro-cell launch), [1] the micro-cell launch, [2] the macro
 * nodes. */
int tsr_solver_time_gfunc_parts(tsr_solver* s, double* g_dev, const double* u_dev, int reps, double* ms3);
/* As tsr_solver_time_gfunc_parts for one k_micro mode: 0 G, 1 G recording the
 * linearisation point, 2 an FD product with the linearised constitutive law (after a
 * mode-1 evaluation at u_dev).  Returns 1 on an unknown mode. */
int tsr_solver_time_gfunc_mode(tsr_solver* s, double* g_dev, const double* u_dev, int mode, int reps, double* ms3);
/* Average device time (ms) of a plain copy of N doubles (reads 8N, writes 8N bytes):
 * the calibration kernel for HBM traffic counters. */
int tsr_solver_time_copy(tsr_solver* s, int reps, double* ms);

/* Sampled per-kernel device timing while the solver steps: HIP events on the solver's
 * stream around every `every`-th launch of each kernel class below (0 turns it off);
 * counters are reset.  _read gives, per class, the launches seen, the launches timed,
 * their summed device time (ms) and summed algorithmic HBM bytes (every scenario
 * active).  Returns 0 (1 on every < 0). */
enum { TSR_KT_MICRO_FULL = 0,   /* k_micro, G or FD product of the full RHS */
       TSR_KT_MICRO_LIN = 1,    /* k_micro, linearised FD product */
       TSR_KT_MACRO_NODE = 2,   /* k_macro_node */
       TSR_KT_DCA = 3,          /* k_pass_dca, DCGS2 projections */
       TSR_KT_DCB = 4,          /* k_pass_dcb, DCGS2 update */
       TSR_KT_PASS = 5,         /* k_pass, CGS2 passes of the plain path */
       TSR_KT_COUNT = 6 };
int tsr_solver_ktimer(tsr_solver* s, int every);
int tsr_solver_ktimer_read(tsr_solver* s, long long* seen, long long* timed, double* ms, double* bytes);

/* Host threads for the per-scenario dense work of a scenario batch (the phi functions
 * of the Krylov stopping rule, the step and the error estimate; scenario q's work
 * touches only scenario q, so results do not depend on the thread count).  The pool has
 * TSR_HOST_THREADS threads if that is set when it is first used, else the hardware
 * threads (at most 16).  n = 1: every scenario on the stepping thread; n = 0 or n > 1:
 * the whole pool; n < 0: query.  Returns the threads a batch now uses.  No counterpart
 * in the reference (its driver steps one problem). */
int tsr_host_threads(int n);

/* Step size the next step starts from (the CurrentStep of exprem_stats_struct, used when
 * a state from elsewhere is resumed).  Returns 1 unless tau is positive and finite. */
int tsr_solver_set_step(tsr_solver* s, double tau);

/* One accepted step on the device state (the exprem contract).  rec may be NULL. */
int tsr_solver_step(tsr_solver* s, double tend, tsr_step_record* rec);
/* Steps until t >= tend or max_steps; records (if non-NULL) receives up to max_records. */
int tsr_solver_run(tsr_solver* s, double tend, int max_steps, tsr_run_stats* out, tsr_step_record* records,
                   int max_records);

/* Step controller: 0 classical (EXPREM's law), 1 learned (MLP proposal, classical fallback),
 * 2 exploration (random walk keeping the Krylov dimension in [3, 24]; generates the
 * learned controller's training data).  Returns 1 on an unknown mode.
 * tsr_solver_set_mlp loads FP32 master weights (layout of tsr_mlp_param_count) and the
 * feature standardisation; tsr_solver_set_seed seeds the exploration controller. */
int tsr_solver_set_controller(tsr_solver* s, int mode);
void tsr_solver_set_seed(tsr_solver* s, unsigned long long seed);
int tsr_solver_set_mlp(tsr_solver* s, const float* params, const float* feat_mean, const float* feat_std);

/* The learned controller's network, 16 -> 64 -> 64 -> 2 with GELU; BF16 weights (round
 * to nearest even of FP32 masters), FP32 biases, activations and accumulation.  No
 * counterpart in the reference (its step law is src/set_exprem_options.cpp:20-26).
 * Parameters: W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] (FP32, row-major).
 * Inputs of forward/grad/train are standardised features x[n][16]; targets t[n][2].
 * Loss of grad: scale * sum_n (w0 (y_n0 - t_n0)^2 + w1 (y_n1 - t_n1)^2).  All return 0,
 * or 2 on a device error (train: 1 on n <= 0 or batch <= 0). */
int   tsr_mlp_param_count(void);
void* tsr_mlp_create(void);
void  tsr_mlp_free(void* mlp);
int   tsr_mlp_set_params(void* mlp, const float* params);
int   tsr_mlp_get_params(void* mlp, float* params);
int   tsr_mlp_forward(void* mlp, const float* x, int n, float* out);
int   tsr_mlp_grad(void* mlp, const float* x, const float* t, int n, float w0, float w1, float scale, float* grad,
                   float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the gradient of the last grad call. */
int   tsr_mlp_adam(void* mlp, float lr);
/* Mini-batch Adam over a seeded permutation per epoch; loss_hist[e] = mean sample loss. */
int   tsr_mlp_train(void* mlp, const float* x, const float* t, int n, int epochs, int batch, float lr, float w0,
                    float w1, unsigned long long seed, float* loss_hist);

/* Solver path.  TSR_PATH_PLAIN (default) is the reference's EXPREM step for step: FD
 * products of the full RHS, CGS2, the Krylov stopping rule checked on the host after
 * every Arnoldi iteration.  The bits of the batched path:
 *   TSR_PATH_DCGS2     Gram-Schmidt with the reorthogonalisation delayed one iteration
 *                      (two passes over the basis per iteration instead of three)
 *   TSR_PATH_DEFERRED  Arnoldi iterations enqueued back to back up to the predicted
 *                      Krylov dimension; the stopping rule is then evaluated for every
 *                      dimension of the window in order (same first passing m)
 *   TSR_PATH_MONOTONE  with DEFERRED: evaluate the stopping rule at the window's last
 *                      dimension first and scan back only when it passes (the rule is
 *                      then satisfied from the chosen m to the window's end; equals the
 *                      first passing m when the rule's estimate decreases with m)
 *   TSR_PATH_JVLIN     FD products with the micro nodes' capacities and conductivities
 *                      expanded to first order about the step's u (the FD quotient then
 *                      differs from the full one by O(eps) terms, as the FD quotient
 *                      itself differs from J v)
 *   TSR_PATH_F32BASIS  with DCGS2: the Gram-Schmidt passes and the step combinations read
 *                      FP32 copies of the Krylov basis columns (FP64 arithmetic; the
 *                      newest column and every FD product stay FP64), halving the
 *                      basis traffic; the basis is then orthonormal to FP32 rounding
 *   TSR_PATH_GRAPH     with DEFERRED: each Arnoldi window (every launch of its iterations
 *                      and of its readback) is captured into a HIP graph on first use and
 *                      replayed after, keyed by basis, window bounds, path and state
 *                      buffer; the same kernels with the same arguments, so the results
 *                      are those of the path without it, bit for bit.  Windows of a host
 *                      RHS (exprem_c with a host Gfunc) and before the micro-cell kernel is
 *                      tuned run uncaptured.  With the kernel timer on (tsr_solver_ktimer),
 *                      each window shape also gets a capture carrying the timer's event
 *                      pairs, replayed every 8th time; turning the timer on or off drops
 *                      the captured graphs.
 * Returns 0, 1 on unknown bits (or F32BASIS without DCGS2, GRAPH without DEFERRED), 2 on
 * allocation failure. */
enum { TSR_PATH_PLAIN = 0, TSR_PATH_DCGS2 = 1, TSR_PATH_DEFERRED = 2, TSR_PATH_JVLIN = 4, TSR_PATH_MONOTONE = 8,
       TSR_PATH_BATCHED = 15, TSR_PATH_F32BASIS = 16, TSR_PATH_GRAPH = 32 };

/* Windows replayed from captured graphs and windows captured since the solver was made
 * (TSR_PATH_GRAPH). */
void tsr_solver_graph_stats(const tsr_solver* s, long long* launches, long long* captures);
int tsr_solver_set_path(tsr_solver* s, int path);
/* jv = (G(u + eps v) - G(u)) / eps about the solver's state u, with the full RHS
 * (lin = 0) or the TSR_PATH_JVLIN expansion (lin = 1); v, jv host arrays of N. */
int tsr_solver_jv(tsr_solver* s, const double* v, double eps, int lin, double* jv);

/* tsr_solver_jv without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and result as tsr_solver_jv; the
 * per-scenario FD step lives in a static device array of 2 x 4096 doubles (one per GPU,
 * kept until the process ends, shared by every call on that GPU: calls that overlap on
 * one GPU race), so the problem may hold at most 4096 scenarios.  lin = 1 needs the
 * linearisation record allocated beforehand (tsr_solver_set_path with TSR_PATH_JVLIN).
 * A micro-cell kernel not tuned yet runs with a fixed launch shape.  Returns 0, 1 on
 * a rejected argument (too many scenarios, no record, eps zero or not finite), 2 on a
 * device error. */
int tsr_solver_jv_noalloc(tsr_solver* s, const double* v, double eps, int lin, double* jv);

/* Raw device pointers (for tests and the bench): u, a scratch vector of length N. */
double* tsr_solver_dev_u(tsr_solver* s);
double* tsr_solver_dev_scratch(tsr_solver* s, int which);
long long tsr_solver_N(const tsr_solver* s);
int tsr_device_sync(void);
/* Device copy helpers for host harnesses. */
int tsr_h2d(double* dst_dev, const double* src_host, long long n);
int tsr_d2h(double* dst_host, const double* src_dev, long long n);

#ifdef __cplusplus
}

/* C++-linkage reference signatures (include/main.h:23-32), for callers compiled
 * against the reference's main.h. */
int macro_mesh_properties(const char* mesh_file, macro_mesh_struct& macro_mesh, int* macro_bc);
int micro_mesh_properties(const char* mesh_file, micro_mesh_struct& micro_mesh);
int read_effective_conductivity(const char* Keff_file, Keff_struct& effective_conductivity);
int set_exprem_options(exprem_options_struct& options);
int exprem(int (*Gfunc)(double* g, double* u, user_data_struct user_data), int N, double& t, double tend,
           double* tspan, double* tspan_solns, double* u0, user_data_struct user_data, exprem_options_struct options,
           exprem_stats_struct& stats);
#endif

#endif /* TSR_H */
