/*
 * bri17_amd -- C ABI of the MI355X (gfx950) matrix-free periodic stiffness
 * operator.  Plain pointers and sizes only; every entry point returns an int
 * status (0 = success) and never throws across the ABI.  The message of the
 * last failure on the calling thread is available from bri17_last_error().
 *
 * Layout of every field: C order (DIM, N0, ..., N_{DIM-1}), component-major,
 * node index row-major with the last axis fastest -- the layout of the
 * reference's u / Ku buffers (tests/test_bri17.cpp:66-67,81-83;
 * include/bri17/bri17.hpp:78-93).  Interleaved complex buffers are
 * (re, im) pairs, the layout of std::complex<T> arrays.
 *
 * Citations are to sbrisard/bri17 v0.1.
 */
#ifndef BRI17_AMD_H
#define BRI17_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRI17_OK 0
#define BRI17_ERR_INVALID 1     /* bad argument (reference: ValueError in pybri17.cpp:13) */
#define BRI17_ERR_UNSUPPORTED 2 /* e.g. a grid row that does not fit the LDS */
#define BRI17_ERR_HIP 3         /* HIP runtime failure (no device, launch error) */
#define BRI17_ERR_NOMEM 4       /* device allocation failed / budget exceeded */
#define BRI17_ERR_STATE 5       /* call not valid in the object's current state */

#define BRI17_F32 0
#define BRI17_F64 1

/* Message of the last failed call on this thread ("" if none). */
const char* bri17_last_error(void);
/* Library version string. */
const char* bri17_version(void);

/* ---------------------------------------------------------------- grid ---- */

/* CartesianGrid::get_node_at (include/bri17/bri17.hpp:78-93); idx has DIM
 * entries. */
int bri17_grid_node_at(int dim, const int* shape, const int* idx, int* node);

/* CartesianGrid::get_cell_nodes (include/bri17/bri17.hpp:127-158): the 2^DIM
 * periodic corners of `cell` in the reference's binary order. */
int bri17_grid_cell_nodes(int dim, const int* shape, int cell, int* nodes);

/* ----------------------------------------------------- Hooke, per frequency */

/* Hooke::modal_strain_displacement (include/bri17/bri17.hpp:212-236).
 * k: DIM ints; B: DIM interleaved complex (2*DIM values). */
int bri17_modal_strain_displacement_f64(int dim, const int* shape, const double* L,
                                        const int* k, double* B);
int bri17_modal_strain_displacement_f32(int dim, const int* shape, const float* L,
                                        const int* k, float* B);

/* Hooke::modal_stiffness (include/bri17/bri17.hpp:247-292).
 * K: DIM*DIM interleaved complex, row-major (imaginary parts are 0). */
int bri17_modal_stiffness_f64(int dim, const int* shape, const double* L, double mu,
                              double nu, const int* k, double* K);
int bri17_modal_stiffness_f32(int dim, const int* shape, const float* L, float mu,
                              float nu, const int* k, float* K);

/* Hooke::modal_eigenstress_to_opposite_strain (include/bri17/bri17.hpp:308-355).
 * tau, eta: SYM = DIM(DIM+1)/2 interleaved complex, Voigt-Mandel order
 * [xx,yy,sqrt2 xy] (2D) / [xx,yy,zz,sqrt2 yz,sqrt2 zx,sqrt2 xy] (3D). */
int bri17_modal_eigenstress_to_opposite_strain_f64(int dim, const int* shape,
                                                   const double* L, double mu,
                                                   double nu, const int* k,
                                                   const double* tau, double* eta);
int bri17_modal_eigenstress_to_opposite_strain_f32(int dim, const int* shape,
                                                   const float* L, float mu, float nu,
                                                   const int* k, const float* tau,
                                                   float* eta);

/* ------------------------------------------------------ operator (device) */

typedef struct bri17_operator bri17_operator;

/* Create the operator of a homogeneous periodic grid (the state held by
 * Hooke{mu, nu, grid}, bri17.hpp:193-194, plus the FFT plans of
 * StiffnessMatrixFactory, tests/test_bri17.cpp:110-128).  Builds the
 * device-resident twiddle and modal tables.  dtype: BRI17_F32 / BRI17_F64.
 * Any grid side, as with the reference's FFTW plans: prime factors up to 13
 * run as butterflies, larger ones as direct-DFT passes (O(p) operations per
 * entry).  Refused with BRI17_ERR_UNSUPPORTED: a transform that needs more
 * than 1,024 threads per sequence or a row beyond the LDS of a workgroup. */
int bri17_operator_create(int dim, const int* shape, const double* L, double mu,
                          double nu, int dtype, bri17_operator** out);
int bri17_operator_destroy(bri17_operator* op);

/* Launch-shape choices of a plan (no reference counterpart: FFTW chooses its
 * own plan inside fftw_plan_dft, tests/test_bri17.cpp:119-127).  Zero fields
 * mean "automatic". */
typedef struct bri17_tuning {
  int plane_split;    /* 3D: split P of the middle axis between the plane and
                         pencil kernels (a divisor of N1, at most 64, with
                         any factorisation) */
  int pencil_width;   /* contiguous spectrum entries per pencil tile */
  int lds_budget_kb;  /* LDS per workgroup, KiB (<= 160) */
  int xcd_remap;      /* 0: off, 1: on, -1: automatic (off for the fixed-size
                         kernels, on for the runtime-size ones) */
  int radices[3][8];  /* radix order of the row / middle-axis / axis-0
                         transforms, 0-terminated (radices 2,3,4,5,7,8,9,16) */
  int force_generic;  /* 1: use the runtime-size kernels even where a
                         fixed-size instantiation exists */
  int plane_prefetch; /* fixed-size kernels: bits 0 / 1 run the forward /
                         inverse plane kernel persistently with the next work
                         item's loads in flight; bit 3 selects the
                         all-components-in-LDS pencil instead of the
                         register-block one (N0 = R^2 grids), bit 2 its
                         persistent variant; -1 = automatic */
} bri17_tuning;

/* bri17_operator_create with explicit tuning (NULL = automatic). */
int bri17_operator_create_tuned(int dim, const int* shape, const double* L, double mu,
                                double nu, int dtype, const bri17_tuning* tuning,
                                bri17_operator** out);

/* JSON description of the operator's plan (kernel geometry, radix orders,
 * LDS sizes) written to buf (truncated to len-1 bytes, NUL-terminated). */
int bri17_operator_describe(const bri17_operator* op, char* buf, size_t len);

/* Change the material of an existing operator (tables are mu/nu independent). */
int bri17_operator_set_material(bri17_operator* op, double mu, double nu);

/* Number of real values of one field (DIM * |N|) and bytes of the operator's
 * half-spectrum workspace. */
int bri17_operator_info(const bri17_operator* op, int64_t* num_dofs,
                        size_t* workspace_bytes);

/* Ku = K u: the apply of StiffnessMatrixFactory::compute_Ku
 * (tests/test_bri17.cpp:56-107) -- DFT, K^(k) u^(k), inverse DFT, times
 * |h|/|N| -- on device pointers, asynchronously on `stream` (a hipStream_t,
 * NULL = default stream).  u and Ku are DIM*|N| values of the operator's
 * dtype; they may not alias.
 *
 * Ordering: an operator owns one half-spectrum workspace, which every device
 * call of that operator uses (apply, solve, strain, eigenstress, its graphs,
 * and batches it is a member of).  Calls of one operator must therefore be
 * ordered on the device -- issued on one stream, or on streams the caller
 * orders with events.  Two operators of the same grid (sharing tables) run
 * concurrently without restriction. */
int bri17_operator_apply(bri17_operator* op, const void* u, void* Ku, void* stream);

/* Green solve K u = f with the reference conventions (no reference entry
 * point; per-k solve as bri17.hpp:340-341, k = 0 mode set to zero as
 * sphinx/theory.rst:208-212 and bri17.hpp:336-339): u^_k = K^_k^-1 f^_k/|h|,
 * u^_0 = 0.  Device pointers, asynchronous on `stream`. */
int bri17_operator_solve(bri17_operator* op, const void* f, void* u, void* stream);

/* Host-pointer variants: copy in, run on the device, copy out, synchronise. */
int bri17_operator_apply_host(bri17_operator* op, const void* u, void* Ku);
int bri17_operator_solve_host(bri17_operator* op, const void* f, void* u);

/* Strain of a displacement field, eps = B u: the operator of the reference's
 * StrainDisplacementMatrixFactory::compute_Bu (tests/test_bri17.cpp:194-238)
 * -- DFT of u, eps^(k) = sym(B^(k) u^(k)) with B^ from
 * Hooke::modal_strain_displacement (bri17.hpp:212-236), inverse DFT / |N|.
 * u: (DIM, N...) nodal field; eps: (SYM, N...) cell field, SYM = 3 (2D) or 6
 * (3D), Voigt-Mandel order [xx, yy, sqrt2 xy] / [xx, yy, zz, sqrt2 yz,
 * sqrt2 zx, sqrt2 xy].  Device pointers, asynchronous on `stream`. */
int bri17_operator_strain(bri17_operator* op, const void* u, void* eps, void* stream);

/* Opposite of the strain induced by an eigenstress field: DFT of tau,
 * Hooke::modal_eigenstress_to_opposite_strain (bri17.hpp:308-355) at every
 * frequency (the per-k loop of python/demo.py:24-40), inverse DFT / |N|.
 * tau, eta: (SYM, N...) cell fields, Voigt-Mandel order.  Device pointers,
 * asynchronous on `stream`.  The first strain / eigenstress call grows the
 * operator's workspace to SYM components (not possible once the operator has
 * been captured in a graph). */
int bri17_operator_eigenstress(bri17_operator* op, const void* tau, void* eta, void* stream);

/* Host-pointer variants (synchronous). */
int bri17_operator_strain_host(bri17_operator* op, const void* u, void* eps);
int bri17_operator_eigenstress_host(bri17_operator* op, const void* tau, void* eta);

/* -------------------------------------------------- no-allocation call */

/* No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The operator call of `mode` -- 0 K u,
 * 1 Green solve, 2 strain, 3 eigenstress, exactly as bri17_operator_apply /
 * _solve / _strain / _eigenstress, on device pointers, asynchronously on
 * `stream` -- with its half-spectrum workspace taken from a static device
 * array of BRI17_NOALLOC_SCRATCH_BYTES that the library holds on every
 * device it runs this on, until the process ends.  It allocates nothing (the
 * operator, created beforehand, holds the tables) and leaves the operator's
 * own workspace untouched.  A grid and mode whose workspace exceeds the
 * array (e.g. the 201 MB of a 256^3 fp32 K u; 128^3 fits in both dtypes) is
 * refused with BRI17_ERR_NOMEM before anything is launched.  All such calls
 * on one device share that device's array: calls that overlap on one device
 * (on different streams, or with a graph replay of such a call) race and
 * give wrong results, so order them. */
#define BRI17_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int bri17_operator_run_noalloc(bri17_operator* op, int mode, const void* in, void* out,
                               void* stream);

/* ---------------------------------------------------------------- batches */

/* Many operators -- different grids, dimensions and materials, one dtype --
 * applied or solved in ONE persistent kernel launch (no reference
 * counterpart: the reference applies one operator per call on the CPU,
 * tests/test_bri17.cpp:56-107).  A batch context owns the device work queue;
 * reuse it across calls. */
typedef struct bri17_batch bri17_batch;
typedef struct bri17_batch_item {
  bri17_operator* op; /* each operator at most once per batch */
  int mode;           /* 0: K u, 1: Green solve */
  const void* in;     /* device field, DIM*|N| values */
  void* out;          /* device field, must not alias in */
} bri17_batch_item;

int bri17_batch_create(bri17_batch** out);
int bri17_batch_destroy(bri17_batch* b);
/* Enqueue the batch on `stream` (asynchronous).  A run whose items equal
 * the previous run's on the same stream (operators, modes, buffers) reuses
 * the device work queue and is enqueued without waiting; any other run first
 * waits for the previous one, then uploads a new queue. */
int bri17_batch_run(bri17_batch* b, const bri17_batch_item* items, int n, void* stream);
/* Wait for the last run; *dependency_timeouts != 0 if a work item gave up
 * waiting for its producers (spin limit reached: results are invalid). */
int bri17_batch_wait(bri17_batch* b, int* dependency_timeouts);
/* Bound on the wait loops of the persistent kernel (default 2^24 polls). */
int bri17_batch_set_spin_limit(bri17_batch* b, int limit);
/* Geometry of the last run. */
int bri17_batch_info(const bri17_batch* b, int* work_items, int* workgroups, int* threads,
                     size_t* lds_bytes);

/* ----------------------------------------------------------------- graphs */

/* A captured HIP graph of one operator call on fixed device buffers, or of
 * one batch, replayed with a single launch (no reference counterpart).  The
 * operator, batch context and buffers must outlive the graph; replays reuse
 * the operator's workspace, so replays of one graph must not overlap.  While
 * a graph lives, the workspaces it replays are pinned (a call that would grow
 * one -- a first strain / eigenstress call -- fails with BRI17_ERR_STATE), and
 * a batch context it was captured from refuses bri17_batch_run and
 * bri17_batch_capture with BRI17_ERR_STATE (its device tables are the
 * graph's); bri17_graph_destroy releases both.
 * mode: 0 apply, 1 solve, 2 strain, 3 eigenstress. */
typedef struct bri17_graph bri17_graph;
/* Runs the call once eagerly, then captures it. */
int bri17_graph_create(bri17_operator* op, int mode, const void* in, void* out,
                       bri17_graph** graph);
/* Runs the batch once eagerly (uploading its work queue), then captures the
 * queue reset and the persistent kernel. */
int bri17_batch_capture(bri17_batch* b, const bri17_batch_item* items, int n,
                        bri17_graph** graph);
int bri17_graph_launch(bri17_graph* graph, void* stream);
int bri17_graph_destroy(bri17_graph* graph);

/* ------------------------------------------------------------- autotuner */

/* Time candidate launch shapes of the operator of a grid (fixed-size kernels
 * and their prefetch variants; runtime-size kernels over the middle-axis
 * split, pencil tile width and radix orders) with `reps` applies each, and
 * return the fastest.  Nothing is registered. */
int bri17_autotune(int dim, const int* shape, const double* L, int dtype, int reps,
                   bri17_tuning* best, double* best_ms, int* candidates_tried);
/* Register the tuning that bri17_operator_create uses for a (dim, shape,
 * dtype); the JSON cache file of the Python package is loaded through it. */
int bri17_tuning_cache_set(int dim, const int* shape, int dtype, const bri17_tuning* t);
int bri17_tuning_cache_clear(void);

/* Per-kernel timing (no reference counterpart; measurement support).
 * While enabled, every apply/solve records HIP events on its stream around
 * its three kernels (R2C plane sweep, fused axis-0 pencil, C2R plane sweep);
 * profile_read waits for them and returns the summed milliseconds of each
 * kernel over `calls` calls, then clears the record. */
int bri17_operator_profile(bri17_operator* op, int enable);
int bri17_operator_profile_read(bri17_operator* op, int* calls, double* ms3);

/* --------------------------------------------------------------- serving
 * No reference counterpart (the reference applies one operator per call,
 * tests/test_bri17.cpp:56-107).  A request loop on one device: requests of
 * mixed grids, dtypes, materials and modes are submitted from any thread; a
 * worker thread admits them in arrival order against an HBM budget and
 * launches them in waves, each wave on its own stream, while earlier waves
 * still run.  Heavy requests (grids of at least `heavy_nodes` nodes) run at
 * most `max_heavy_waves` at a time; a heavy request that has to wait for that
 * lets the requests behind it start, so small grids are not held behind
 * several large ones.  A wave is either one persistent batched launch of requests
 * with runtime-size plans (bri17_batch_*), or one request that runs alone on
 * its plan's own kernels (grids with fixed-size kernels, and grids of at
 * least `solo_nodes` nodes).  Which of the two a request takes depends on the
 * request alone, so its result does not depend on what it was served with.
 * Plans (twiddle and modal tables) stay resident per (grid, dtype).
 *
 * HBM footprint of a request = its input + output fields + its half-spectrum
 * workspace.  Admission holds the sum over admitted requests, plus the idle
 * workspaces and staging buffers the server caches, within the budget; a
 * request whose footprint alone exceeds the budget is rejected at submit.  */

typedef struct bri17_server bri17_server;

typedef struct bri17_server_config {
  size_t hbm_budget_bytes; /* 0: half of the free device memory at creation */
  int max_batch;           /* requests per batched wave (0: 64) */
  int max_waves;           /* waves in flight, one stream each (0: 4, max 16) */
  int64_t solo_nodes;      /* grids of at least this many nodes run alone
                              (0: 2^21, -1: never) */
  int spin_limit;          /* dependency-wait bound of batched waves (0: default) */
  int64_t heavy_nodes;     /* grids of at least this many nodes are heavy (0: 2^22) */
  int max_heavy_waves;     /* heavy waves in flight at once (0: 1, -1: no limit);
                              while the limit is reached, a queued heavy request
                              lets the requests behind it start first */
} bri17_server_config;

#define BRI17_MEM_DEVICE 0 /* in / out are device pointers owned by the caller */
#define BRI17_MEM_HOST 1   /* in / out are host pointers; the server stages them */

typedef struct bri17_request {
  int dim;          /* 2 or 3 */
  int shape[3];     /* N0, N1[, N2] */
  double L[3];      /* cell sizes of the periodic box (L0, L1[, L2]) */
  double mu, nu;    /* shear modulus and Poisson ratio of this request */
  int mode;         /* 0: K u, 1: Green solve (as bri17_operator_apply / solve) */
  int dtype;        /* BRI17_F32 or BRI17_F64 */
  int memory;       /* BRI17_MEM_DEVICE or BRI17_MEM_HOST */
  const void* in;   /* field (DIM, N0, ...); must stay valid until completion */
  void* out;
} bri17_request;

#define BRI17_REQ_QUEUED 0
#define BRI17_REQ_RUNNING 1
#define BRI17_REQ_DONE 2
#define BRI17_REQ_FAILED 3
#define BRI17_REQ_CANCELLED 4

typedef struct bri17_request_result {
  int state;          /* BRI17_REQ_* */
  int error;          /* status code of a failed request (0 otherwise) */
  double queued_s;    /* submit -> launch */
  double latency_s;   /* submit -> completion observed by the server */
  int64_t wave;       /* wave the request ran in */
  int wave_size;      /* requests in that wave */
} bri17_request_result;

typedef struct bri17_server_stats {
  int64_t submitted, completed, failed, rejected, cancelled;
  int64_t waves, batched_waves, solo_waves;
  int64_t queued, running;
  size_t hbm_budget, hbm_admitted, hbm_cached, hbm_peak;
  int64_t plans_resident;
  int64_t peak_waves_in_flight; /* most waves seen on the device at once */
} bri17_server_stats;

int bri17_server_create(const bri17_server_config* cfg, bri17_server** server);
/* Waits for every admitted request, then frees everything. */
int bri17_server_destroy(bri17_server* server);
/* Queue a request; returns its id.  Fails at once (nothing queued) on an
 * invalid request, an unsupported grid, or a footprint above the budget. */
int bri17_server_submit(bri17_server* server, const bri17_request* req, int64_t* id);
/* Queue n requests at once: they reach the worker together, so they can
 * share waves.  All or nothing: if one is rejected, none is queued. */
int bri17_server_submit_many(bri17_server* server, const bri17_request* reqs, int n,
                             int64_t* ids);
/* Non-blocking state of a request.  A request in a final state (done,
 * failed, cancelled) is forgotten once its result has been returned. */
int bri17_server_poll(bri17_server* server, int64_t id, bri17_request_result* res);
/* Block until the request reaches a final state or `timeout_s` elapses
 * (negative: no limit); BRI17_ERR_STATE on timeout. */
int bri17_server_wait(bri17_server* server, int64_t id, double timeout_s,
                      bri17_request_result* res);
/* Cancel a request that is still queued (BRI17_ERR_STATE once launched). */
int bri17_server_cancel(bri17_server* server, int64_t id);
/* Block until nothing is queued or running. */
int bri17_server_drain(bri17_server* server);
int bri17_server_get_stats(bri17_server* server, bri17_server_stats* stats);
/* Release the idle cached workspaces and staging buffers. */
int bri17_server_trim(bri17_server* server);
/* Dependency-wait bound of later batched waves (a wave whose wait runs out
 * fails with BRI17_ERR_STATE; the server goes on with the next waves). */
int bri17_server_set_spin_limit(bri17_server* server, int limit);

#ifdef __cplusplus
}
#endif
#endif
