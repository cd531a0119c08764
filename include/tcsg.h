/*
 * tcsg.h -- C ABI of the MI355X thread-coarsened stencil runtime (libtcsg.so).
 *
 * This is the drop-in boundary for the reference's generated launch
 * interface.  The reference's host harness reaches its kernels through four
 * C++ functions declared in hip/launchUpdateGrid.h:13-18 and emitted by
 * hip/generateLaunchUpdateGrid.py; each entry point below names the one it
 * replaces.  Differences that are deliberate:
 *   - plain C linkage, status codes instead of exit()/throw, no hip headers
 *     (streams and device pointers are void*);
 *   - a loaded kernel is a handle (tcsg_kernel) owning its code object and its
 *     mask buffer, so several stencils can live in one process (the
 *     reference keeps one __constant__ mask per process);
 *   - one-step launches take an explicit stream and do not synchronise.
 *
 * Kernels are code objects written by the generator
 * (threadcoarsenedstencilgenerator_amd/generator) and compiled for gfx950.
 */
#ifndef TCSG_H
#define TCSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enumerations (values match the generator's spec module) ---- */
enum { TCSG_BOX = 0, TCSG_STAR = 1 };
enum { TCSG_ROW = 0, TCSG_COLUMN = 1 };
enum { TCSG_JACOBI = 0, TCSG_VARIABLE = 1 };
enum { TCSG_ALIGN_NONE = 0, TCSG_ALIGN_READ = 1, TCSG_ALIGN_WRITE = 2 };
enum { TCSG_F64 = 0, TCSG_F32C = 1 };
enum { TCSG_PLAIN = 0, TCSG_STREAM = 1, TCSG_TEMPORAL = 2 };

/* initial-value modes for tcsg_launch_init */
enum {
    TCSG_INIT_REFERENCE = 0, /* the reference's init kernel values        */
    TCSG_INIT_U01 = 1,       /* interior U[0,1), halo 0 (SplitMix64)       */
    TCSG_INIT_UPM1 = 2,      /* interior U[-1,1), halo 0                   */
    TCSG_INIT_NORMAL = 3     /* interior N(0,1) by Box-Muller, halo 0      */
};

/* ---- status codes ---- */
enum {
    TCSG_OK = 0,
    TCSG_ERR_ARG = 1,          /* invalid argument                                   */
    TCSG_ERR_HIP = 2,          /* a HIP runtime call failed (see tcsg_last_hip_error) */
    TCSG_ERR_MISMATCH = 3,     /* code object parameters differ from the descriptor  */
    TCSG_ERR_BLOCK_SHAPE = 4,  /* block shape not launchable for this kernel         */
    TCSG_ERR_TOO_LARGE = 5,    /* grid does not fit this kernel's addressing         */
    TCSG_ERR_NO_MASK = 6,      /* Variable kernel launched before tcsg_set_mask      */
    TCSG_ERR_FILE = 7          /* code object missing or unreadable                  */
};

typedef struct tcsg_dim3 { uint32_t x, y, z; } tcsg_dim3;

/* The generator's compile-time parameters (the reference generator argv,
 * hip/generate2DStencil_hip.py:28-44, hip/generate3DStencil_hip.py:29-46,
 * plus the element type and code shape). */
typedef struct tcsg_desc {
    uint32_t dimension;      /* 2 or 3                      */
    uint32_t neighbourhood;  /* TCSG_BOX / TCSG_STAR        */
    uint32_t orientation;    /* TCSG_ROW / TCSG_COLUMN      */
    uint32_t coeff;          /* TCSG_JACOBI / TCSG_VARIABLE */
    uint32_t range;          /* r >= 1                      */
    uint32_t xc, yc, zc;     /* coarsening (zc = 1 in 2D)   */
    uint32_t alignment;      /* TCSG_ALIGN_*                */
    uint32_t cl_size;        /* alignment size in points    */
    uint32_t dtype;          /* TCSG_F64 / TCSG_F32C        */
    uint32_t variant;        /* TCSG_PLAIN / TCSG_STREAM / TCSG_TEMPORAL */
} tcsg_desc;

/* Padded geometry of a grid (hip/benchmark.cpp:140-177, hip/benchmark.h:100-123). */
typedef struct tcsg_geometry {
    tcsg_dim3 size;       /* allocation extents incl. halo and padding (the reference's gridSize) */
    tcsg_dim3 elements;   /* interior points per axis (gridElements)  */
    tcsg_dim3 threads;    /* threads per axis (gridThreads)           */
    uint32_t left_offset; /* padding before each row (Write alignment) */
    uint32_t right_offset;/* padding after each row                    */
    uint32_t dimension;
    uint32_t range;
    uint64_t n_points;    /* size.x * size.y * size.z (64-bit)         */
} tcsg_geometry;

typedef struct tcsg_kernel tcsg_kernel;

/* Geometry for a user grid size incl. halo (`size`, z ignored in 2D).
 * Replaces the geometry part of stencilBenchmark::initBenchmark
 * (hip/benchmark.cpp:133-182). */
int tcsg_geometry_compute(const tcsg_desc *desc, tcsg_dim3 size, tcsg_geometry *out);

/* TCSG_OK if *g is the geometry this descriptor produces for some user grid
 * size and, for TCSG_STREAM kernels, fits their 32-bit buffer addressing;
 * TCSG_ERR_ARG / TCSG_ERR_TOO_LARGE otherwise.  Every launch entry point
 * applies this check. */
int tcsg_geometry_check(const tcsg_desc *desc, const tcsg_geometry *g);

/* Load a generated code object on the current device and check its baked-in
 * parameters against *expect (NULL: accept).  Replaces the compile-time
 * binding of kernels.cu into launchUpdateGrid.cu
 * (hip/generateLaunchUpdateGrid.py:55-58) and its sanity check (:106-119). */
int tcsg_kernel_load(tcsg_kernel **out, const char *code_object_path, const tcsg_desc *expect);
int tcsg_kernel_free(tcsg_kernel *k);
int tcsg_kernel_desc(const tcsg_kernel *k, tcsg_desc *out);

/* Geometry for this kernel: tcsg_geometry_compute, except that a kernel
 * generated with remainder handling (a STREAM option) keeps every point of an
 * axis whose interior is not a multiple of the coarsening; threads per axis
 * are then rounded up; the last thread in x stores only the columns inside,
 * in y and z it starts early enough to end at the interior's end (which needs
 * an interior of at least one coarsening on those axes: TCSG_ERR_ARG otherwise).
 * The reference truncates such axes (hip/benchmark.cpp:139-161), which is
 * what every other kernel does.  The launch entry points accept the geometry
 * this returns. */
int tcsg_kernel_geometry(const tcsg_kernel *k, tcsg_dim3 size, tcsg_geometry *out);
/* *out = 1 for a kernel generated with remainder handling, else 0. */
int tcsg_kernel_remainder(const tcsg_kernel *k, int *out);
/* *out = time steps one launch advances: 1, or for a TCSG_TEMPORAL kernel (2D,
 * no reference counterpart) the steps it runs per launch in LDS; tcsg_update_grid,
 * tcsg_run and tcsg_time_steps then count launches, each of that many steps.
 * A temporal kernel runs with block (64, launch bound / 64, 1) only and keeps
 * every interior point whatever the remainder option. */
int tcsg_kernel_steps_per_launch(const tcsg_kernel *k, int *out);

/* Variable-mask upload (n doubles in the layout of the reference's d_mask).
 * Replaces copyMask (hip/launchUpdateGrid.h:18, hip/generateLaunchUpdateGrid.py:77-82).
 * F32C kernels receive the FP32 rounding of each weight. */
int tcsg_set_mask(tcsg_kernel *k, const double *mask, uint32_t n);

/* Block-size limits.  occupancy_optimal is what the reference's
 * maxPotentialBlockSize returns (hip/launchUpdateGrid.h:17,
 * hip/generateLaunchUpdateGrid.py:84-89); legal_max is the largest block the
 * kernel can be launched with.  Either pointer may be NULL. */
int tcsg_max_block_size(tcsg_kernel *k, int *occupancy_optimal, int *legal_max);

/* 0 if `block` can be launched with this kernel, else TCSG_ERR_BLOCK_SHAPE. */
int tcsg_check_block(const tcsg_kernel *k, tcsg_dim3 block);

/* Resources of the loaded kernel (the block-shape cost model's inputs):
 * VGPRs per lane, private scratch bytes per lane,
 * and, for a block of `threads` threads, the workgroups the occupancy
 * calculator lets one CU hold (0 if none fit).  Any output may be NULL.
 * No reference counterpart: the reference's sweep measures every shape
 * (hip/benchmark.cpp:221-263) and needs no prediction inputs. */
int tcsg_kernel_resources(tcsg_kernel *k, uint32_t threads, int *vgprs, int *scratch_bytes, int *workgroups_per_cu);

/* Blocks per axis = ceil(threads / block) (hip/benchmark.cpp:232,301-303). */
tcsg_dim3 tcsg_grid_blocks(const tcsg_geometry *g, tcsg_dim3 block);

/* Fill a whole grid buffer (elements of the kernel's dtype) with initial
 * values.  TCSG_INIT_REFERENCE replaces launchInit
 * (hip/launchUpdateGrid.h:14, hip/generateLaunchUpdateGrid.py:150-158); the
 * seeded modes are counter-based on the interior index.  Padding is zeroed.
 * Asynchronous on `stream`. */
int tcsg_launch_init(tcsg_kernel *k, const tcsg_geometry *g, int mode, uint64_t seed,
                     void *d_array, void *stream);

/* One time step: out = stencil(in) on the interior (halo and padding of
 * `out` untouched).  The launch in the reference's benchmark loop
 * (hip/generateLaunchUpdateGrid.py:92-100,131).  Asynchronous. */
int tcsg_update_grid(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks,
                     const void *in, void *out, void *stream);

/* `steps` ping-pong steps (step t reads buffer t%2), asynchronous.  *final
 * receives the index (0/1) of the buffer holding the result. */
int tcsg_run(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks,
             void *d0, void *d1, int steps, void *stream, int *final_index);

/* The reference's timing protocol (hip/generateLaunchUpdateGrid.py:103-147):
 * doubling rounds of 1,2,4,... steps with a device synchronisation after
 * every step and a host pointer swap, until the elapsed time exceeds
 * max_time; returns the 2^k-1 per-step durations [s] in durations (up to
 * cap; *n gets the full count).  Null stream.  The rounds count launches: a
 * temporal kernel's launch advances steps_per_launch steps and contributes that
 * many durations, each an equal share of the launch. */
int tcsg_benchmark_update_grid(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block,
                               tcsg_dim3 blocks, double max_time, void *d0, void *d1,
                               double *durations, size_t cap, size_t *n);

/* Event timing of `steps` back-to-back steps after `warmup` untimed ones,
 * on `stream`, no host synchronisation between steps.  *ms = total time. */
int tcsg_time_steps(tcsg_kernel *k, const tcsg_geometry *g, tcsg_dim3 block, tcsg_dim3 blocks,
                    void *d0, void *d1, int warmup, int steps, void *stream, float *ms);

/* ---- batches of independent grids (kernels generated with the batched
 * option; no reference counterpart) ----
 * One persistent kernel per time step advances every grid of the batch: its
 * workgroups take tiles, (grid, block) pairs each the work of one block of
 * tcsg_update, from a work queue over the tile table, one contiguous range per
 * XCD (a workgroup drains its XCD's range, then the others).  Kernels generated
 * with packed tiles take instead runs of block.x * block.y consecutive threads
 * of a grid's xy thread space (x fastest, continuing on the next thread row)
 * times block.z thread planes, so uneven grids leave no lanes idle.  geoms[i] is
 * grid i's geometry, bufs0[i]/bufs1[i] its ping-pong buffers; every tile uses
 * `block`; workgroups = 0 launches as many workgroups per compute unit as the
 * kernel's occupancy allows at that block size (at most one per tile).  The
 * queue is the batch's own device memory, cleared by each launch's last
 * workgroup: runs of one batch must not overlap in time. */
typedef struct tcsg_batch tcsg_batch;
int tcsg_batch_create(tcsg_batch **out, tcsg_kernel *k, uint32_t n, const tcsg_geometry *geoms,
                      void *const *bufs0, void *const *bufs1, tcsg_dim3 block, uint32_t workgroups);
int tcsg_batch_free(tcsg_batch *b);
/* `steps` steps of every grid starting from bufs0 (step t reads buffer t%2),
 * one persistent launch per step.  use_graph != 0: the launches are captured
 * once into a hipGraph (kept per step count) and the graph is replayed.
 * Asynchronous on `stream` (NULL: the library's own stream, synchronised
 * before returning).  *final_index = steps % 2. */
int tcsg_batch_run(tcsg_batch *b, int steps, int use_graph, void *stream, int *final_index);
/* The same steps with every grid launched separately (tcsg_update_grid per
 * grid per step) -- the baseline the persistent kernel replaces. */
int tcsg_batch_run_separate(tcsg_batch *b, int steps, void *stream, int *final_index);
/* Total interior points of the batch and its tile count. */
int tcsg_batch_info(const tcsg_batch *b, uint64_t *interior_points, uint32_t *tiles, uint32_t *workgroups);

/* Error of an F32C state `a` (layout ga) against an FP64 state `b` (layout
 * gb) over their common interior, computed in FP64 on the GPU:
 * out[0] = max |a-b|, out[1] = max |b|, out[2] = sum (a-b)^2,
 * out[3] = sum b^2, out[4] = max |a-b|/max(|b|, floor).  The layouts may
 * differ in padding (FP32 rows aligned to their own 128-byte lines) but must
 * have the same interior.  Synchronous (null stream).  Its 5-double reduction
 * scratch is allocated on the first call on a device and kept; calls that
 * overlap on one device race on it. */
int tcsg_error_f32_vs_f64(const tcsg_geometry *ga, const float *a, const tcsg_geometry *gb, const double *b,
                          double floor, double *out);

/* The same error reduction without allocating anything, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Asynchronous on
 * `stream`; the 5 results (as above) go to the device array `out`.  Two passes with
 * a fixed summation order (deterministic, unlike tcsg_error_f32_vs_f64's atomics):
 * per-workgroup partial results in the library's static device scratch of 16,384
 * rows of 5 doubles (640 KiB on each GPU it runs on, held until the process ends),
 * so interiors of at most 2^30 points; larger ones return TCSG_ERR_TOO_LARGE.
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU race. */
int tcsg_error_f32_vs_f64_noalloc(const tcsg_geometry *ga, const float *a, const tcsg_geometry *gb,
                                  const double *b, double floor_v, double *out, void *stream);

/* FP64 -> FP32 copy of a whole grid (round to nearest).  Asynchronous. */
int tcsg_convert_f64_to_f32(const double *src, float *dst, uint64_t n, void *stream);

/* HBM bandwidth probe: `reps` launches (after one warm-up) of a plain
 * 16-byte-per-lane stream kernel over `bytes` bytes; mode 0 copies src to dst,
 * 1 reads src only, 2 writes dst only; nontemporal selects nontemporal loads
 * and stores; workgroups 0 = one pass without a grid-stride loop.  Synchronous;
 * *ms_per_rep gets the HIP-event time per launch.  The copy rate is the
 * practical ceiling of a one-step stencil (one grid read, one written).
 * No reference counterpart (the reference reports no bandwidth ceiling). */
int tcsg_hbm_probe(int mode, int nontemporal, const void *src, void *dst, uint64_t bytes, uint32_t workgroups,
                   int reps, void *stream, float *ms_per_rep);

/* The same probe with `unroll` 16-byte elements per lane per iteration (1, 2,
 * 4, 8 or 16: that many loads in flight per lane before the stores) and
 * `block` threads per workgroup (256, 512 or 1024); tcsg_hbm_probe is unroll 4,
 * block 256.  No reference counterpart. */
int tcsg_hbm_probe_ex(int mode, int nontemporal, uint32_t unroll, uint32_t block, const void *src, void *dst,
                      uint64_t bytes, uint32_t workgroups, int reps, void *stream, float *ms_per_rep);

/* ---- cost-model network of the learned block-shape tuner (no reference
 * counterpart: the reference measures every block shape, hip/benchmark.cpp:
 * 221-263; the tuner ranks them with this network and measures the top few).
 * A 3-layer MLP  y = W3 relu(W2 relu(W1 x + b1) + b2) + b3  in FP32 with
 * `hidden` units per hidden layer, trained on a mean-squared error with Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).  Parameters are one
 * array in the order W1[hidden][in], b1, W2[hidden][hidden], b2, W3[hidden],
 * b3.  Device pointers; the network owns its parameters and activations. */
typedef struct tcsg_mlp tcsg_mlp;
size_t tcsg_mlp_num_params(uint32_t in_dim, uint32_t hidden);
/* Weights U[-1/sqrt(fan_in), 1/sqrt(fan_in)) drawn from `seed` (SplitMix64), zero biases. */
int tcsg_mlp_create(tcsg_mlp **out, uint32_t in_dim, uint32_t hidden, uint32_t max_batch, uint64_t seed);
int tcsg_mlp_free(tcsg_mlp *n);
/* Host copies of the parameters; set also resets the Adam state. */
int tcsg_mlp_get_params(const tcsg_mlp *n, float *host, size_t count);
int tcsg_mlp_set_params(tcsg_mlp *n, const float *host, size_t count);
/* d_out[b] = network(d_x[b][0..in)) for b < batch.  Asynchronous on stream. */
int tcsg_mlp_forward(tcsg_mlp *n, const float *d_x, uint32_t batch, float *d_out, void *stream);
/* Gradient of the mean-squared error over the batch with respect to every
 * parameter (host copy, same layout) and the loss.  Synchronous; the
 * parameters are unchanged. */
int tcsg_mlp_gradients(tcsg_mlp *n, const float *d_x, const float *d_y, uint32_t batch, float *host_grads,
                       float *host_loss);
/* One Adam step on the batch's mean-squared error with learning rate lr;
 * *d_loss (device, optional) receives the loss before the step.
 * Asynchronous on stream. */
int tcsg_mlp_train_step(tcsg_mlp *n, const float *d_x, const float *d_y, uint32_t batch, float lr, float *d_loss,
                        void *stream);

/* Readable name of a status code; last HIP error code seen by the library. */
const char *tcsg_status_string(int status);
int tcsg_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif /* TCSG_H */
