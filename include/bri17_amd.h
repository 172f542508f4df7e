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
#define BRI17_ERR_UNSUPPORTED 2 /* e.g. a grid side with a prime factor > 7 */
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
 * Grid sides must factor into 2, 3, 5, 7. */
int bri17_operator_create(int dim, const int* shape, const double* L, double mu,
                          double nu, int dtype, bri17_operator** out);
int bri17_operator_destroy(bri17_operator* op);

/* Launch-shape choices of a plan (no reference counterpart: FFTW chooses its
 * own plan inside fftw_plan_dft, tests/test_bri17.cpp:119-127).  Zero fields
 * mean "automatic". */
typedef struct bri17_tuning {
  int plane_split;    /* 3D: split P of the middle axis between the plane and
                         pencil kernels (1,2,3,4,5,7,8,16) */
  int pencil_width;   /* contiguous spectrum entries per pencil tile */
  int lds_budget_kb;  /* LDS per workgroup, KiB (<= 160) */
  int xcd_remap;      /* 0: off, 1: on, -1: automatic (on) */
  int radices[3][8];  /* radix order of the row / middle-axis / axis-0
                         transforms, 0-terminated (radices 2,3,4,5,7,8,16) */
  int force_generic;  /* 1: use the runtime-size kernels even where a
                         fixed-size instantiation exists */
  int plane_prefetch; /* fixed-size plane kernels run persistently with the
                         next plane's loads in flight: bit 0 forward, bit 1
                         inverse; -1 = automatic */
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
 * dtype; they may not alias. */
int bri17_operator_apply(bri17_operator* op, const void* u, void* Ku, void* stream);

/* Green solve K u = f with the reference conventions (no reference entry
 * point; per-k solve as bri17.hpp:340-341, k = 0 mode set to zero as
 * sphinx/theory.rst:208-212 and bri17.hpp:336-339): u^_k = K^_k^-1 f^_k/|h|,
 * u^_0 = 0.  Device pointers, asynchronous on `stream`. */
int bri17_operator_solve(bri17_operator* op, const void* f, void* u, void* stream);

/* Host-pointer variants: copy in, run on the device, copy out, synchronise. */
int bri17_operator_apply_host(bri17_operator* op, const void* u, void* Ku);
int bri17_operator_solve_host(bri17_operator* op, const void* f, void* u);

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
/* Enqueue the batch on `stream` (asynchronous). */
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

/* A captured HIP graph of one apply / solve on fixed device buffers, or of
 * one batch, replayed with a single launch (no reference counterpart).  The
 * operator, batch context and buffers must outlive the graph; replays reuse
 * the operator's workspace, so replays of one graph must not overlap. */
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

#ifdef __cplusplus
}
#endif
#endif
