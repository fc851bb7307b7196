/*
 * pstm_hip.h — C ABI of the MI355X (gfx950) PSTM migration backend.
 *
 * The entry points replace, for the `hip` backend, the device boundary the
 * reference crosses inside MigrationKernel.migrate_tile
 * (pstm/kernels/base.py:290-318): the registered Numba CPU backend calls
 * _migrate_tile_kernel_optimized (pstm/kernels/numba_cpu_optimized.py:93-233)
 * after _precompute_time_terms (:358-401); the Metal precedent crosses it with
 * the pybind11 call pstm_metal.migrate_tile (pstm/metal/src/bindings.cpp:48-159).
 *
 * Conventions (the reference's pybind module raises; here nothing crosses the
 * C ABI but status codes):
 *   - every function returns 0 on success, non-zero on failure; the message of
 *     the last failure on the calling thread is pstm_hip_last_error();
 *   - all pointers in PstmMigrateArgs are DEVICE pointers on the current HIP
 *     device, the stream argument is a hipStream_t (NULL = default stream);
 *   - image and fold are ACCUMULATED (+=), as the Numba backend does
 *     (numba_cpu_optimized.py:229,231);
 *   - launches are asynchronous on `stream`; pstm_hip_stream_sync() waits.
 */
#ifndef PSTM_HIP_H
#define PSTM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSTM_HIP_ABI_VERSION 5

/* traveltime: KernelConfig.kernel_type (pstm/kernels/base.py:220-229) */
#define PSTM_TT_STRAIGHT 0 /* DSR, numba_cpu_optimized.py:189-194            */
#define PSTM_TT_CURVED 1   /* V(z)=V0+kz, pstm/algorithm/curved_ray.py:19-52 */
#define PSTM_TT_VTI 2      /* eta, pstm/algorithm/anisotropy_vti.py:148-220 */
/* VTI with the learned traveltime surrogate evaluated inside the gather (BF16
 * MFMA, f32 accumulation); no counterpart in the reference.  See the mlp_*
 * fields below. */
#define PSTM_TT_VTI_LEARNED 3
/* Curved ray (V(z) = V0 + kz) with the learned traveltime surrogate: the same
 * fused network and t0 lattice, inputs (tau, s_s, s_r, vg, k) with z the
 * depth of t0 / 2, vg = sqrt(v0 v(z)) and s = k h / sqrt(vg^2 + k^2 h^2);
 * outputs out_scale * c per leg, leg time z / sqrt(vg^2 + k^2 h^2) (1 + c)
 * (the reference's log formula equals z / sqrt(xi^2 + k^2 h^2) for some xi
 * between v0 and v(z), so c stays small), trained on the legs of
 * pstm/algorithm/curved_ray.py:19-52, 239-262.  Needs |k| >= 1e-6.  No
 * counterpart in the reference.  mlp_v_lo / mlp_v_hi scale vg, mlp_eta_max
 * scales k. */
#define PSTM_TT_CURVED_LEARNED 4

/* anti-aliasing (the reference's registered kernel has none) */
#define PSTM_AA_NONE 0
#define PSTM_AA_TRIANGLE 1 /* triangle filter on double-integrated traces */

/* trace interpolation: KernelConfig.interpolation_method, honoured by the
 * reference's NumbaKernel through interpolate_sample
 * (pstm/kernels/interpolation.py:276-356, numba_cpu.py:341,632).  The codes
 * are the reference's get_method_code() numbers except that 0 is linear (the
 * default of a zeroed struct) and 1 nearest.  Methods other than linear need
 * the straight-ray traveltime, no anti-aliasing and no E counter, as in the
 * reference, whose NumbaKernel is the only backend that honours them. */
#define PSTM_INTERP_LINEAR 0   /* _linear_interp            :68-92   */
#define PSTM_INTERP_NEAREST 1  /* _nearest_interp           :58-65   */
#define PSTM_INTERP_CUBIC 2    /* Catmull-Rom, _cubic_interp :95-135 */
#define PSTM_INTERP_SINC4 3    /* Hann-windowed sinc, half width 2, :154-203 */
#define PSTM_INTERP_SINC8 4    /* half width 4             */
#define PSTM_INTERP_SINC16 5   /* half width 8             */
#define PSTM_INTERP_LANCZOS3 6 /* _lanczos_interp a = 3, :212-268 */
#define PSTM_INTERP_LANCZOS5 7 /* a = 5                    */

/* velocity slice layout: VelocitySlice.is_1d (pstm/kernels/base.py:180-194) */
#define PSTM_VEL_1D 0 /* vrms (nt,)        */
#define PSTM_VEL_3D 1 /* vrms (nx, ny, nt) */

/* eta layout for PSTM_TT_VTI (KernelConfig.vti_eta_*, base.py:231-235) */
#define PSTM_ETA_CONST 0
#define PSTM_ETA_1D 1
#define PSTM_ETA_3D 2

/*
 * One tile migration: TraceBlock (base.py:27-116) x OutputTile (base.py:
 * 119-177) x VelocitySlice (base.py:180-194) x KernelConfig (base.py:197-265).
 * Plain old data; mirrored field for field by a ctypes.Structure and checked
 * with pstm_hip_args_size() when the library is loaded.
 */
typedef struct PstmMigrateArgs {
    /* --- trace block (device) --- */
    const float* amplitudes;      /* (rows, n_samples) row-major, f32     */
    const double* source_x;       /* (rows,) absolute coordinates, f64    */
    const double* source_y;
    const double* receiver_x;
    const double* receiver_y;
    const double* midpoint_x;
    const double* midpoint_y;
    const double* trace_weights;  /* (rows,) or NULL = all 1.0            */
    const int32_t* trace_index;   /* (n_traces,) rows to use, or NULL     */
    /* --- output tile (device) --- */
    const double* x_axis;         /* (nx,)                                */
    const double* y_axis;         /* (ny,)                                */
    const double* t_axis_ms;      /* (nt,)                                */
    /* --- velocity slice (device) --- */
    const double* vrms;           /* (nt,) or (nx, ny, nt)                */
    const double* eta;            /* NULL, (nt,) or (nx, ny, nt)          */
    /* --- outputs (device, accumulated) --- */
    float* image;                 /* (nx, ny, nt) f32, +=                 */
    int32_t* fold;                /* (nx, ny) pillar fold, +=             */
    unsigned long long* n_effective; /* NULL or one counter, += E         */
    /* --- sizes and switches --- */
    int32_t n_traces;             /* traces used (length of trace_index)  */
    int32_t n_samples;            /* samples per trace                    */
    int32_t nx, ny, nt;
    int32_t vel_mode;             /* PSTM_VEL_*                           */
    int32_t eta_mode;             /* PSTM_ETA_*                           */
    int32_t traveltime;           /* PSTM_TT_*                            */
    int32_t aa_mode;              /* PSTM_AA_*                            */
    int32_t apply_spreading;      /* 1/(v t) weight, t > 1 ms             */
    int32_t apply_obliquity;      /* t0/t weight, t > 1 ms                */
    int32_t aa_max_half_width;    /* triangle half width clamp (samples)  */
    /* --- scalars --- */
    double dt_in_ms;              /* TraceBlock.sample_rate_ms            */
    double t_start_in_ms;         /* TraceBlock.start_time_ms             */
    double max_aperture_m;
    double min_aperture_m;
    double max_dip_deg;
    double taper_fraction;
    double grid_dx;               /* output bin size x (AA operator dip)  */
    double grid_dy;               /* output bin size y                    */
    double curved_v0;             /* PSTM_TT_CURVED: V(z) = v0 + k z      */
    double curved_k;
    double eta_constant;          /* PSTM_ETA_CONST                       */
    /* --- learned traveltime (PSTM_TT_VTI_LEARNED, PSTM_TT_CURVED_LEARNED) ---
     * Per (pillar, trace) pair the network is evaluated on a lattice of output
     * samples n_0 = 0, n_{j+1} = min(n_j + min(max((n_j g) >> 8, K), K_max),
     * nt - 1) up to the last sample nt - 1 (K = mlp_node_spacing, g =
     * mlp_node_growth, K_max = mlp_node_max or K; g = 0: uniform spacing K),
     * only at nodes read by some sample inside the pair's aperture, with
     * inputs
     *   x = (2 tau / tau_max - 1, 2 s_s - 1, 2 s_r - 1,
     *        2 (v - v_lo) / (v_hi - v_lo) - 1, 2 eta / eta_max - 1),
     * tau = t0 / 2, v and eta at (pillar, t0), s = sine of the straight-ray
     * angle of a leg, d / (v sqrt(tau^2 + d^2 / v^2)).  Outputs are
     * out_scale * g for each leg, g = c / eta the anisotropic factor c of the
     * leg, (t^2 - tau^2) / (d / v)^2 - 1, per unit eta (for the exact VTI
     * time g = -2 s^2 / (1 + 2 eta s^2), in [-2, 0]); g is interpolated
     * linearly between lattice nodes and the leg time is
     * sqrt(tau^2 + (d / v)^2 max(1 + eta g, 0)) with v and eta at the output
     * sample itself, t = t0 for t0 <= 0 as in the exact VTI time.  Carrying
     * g rather than c leaves the lattice to follow only the factor's own
     * change with t0, not that of eta in the model.  Needs nt >= 2.        */
    const float* mlp_params;      /* PSTM_MLP_PARAMS f32 (network layout below);
                                     weights are rounded to bf16 on use       */
    int32_t mlp_node_spacing;     /* K: smallest node spacing, 4 .. 512
                                     output samples                           */
    int32_t interpolation;        /* PSTM_INTERP_* (0 = linear)               */
    double mlp_tau_max;
    double mlp_v_lo;
    double mlp_v_hi;
    double mlp_eta_max;
    double mlp_out_scale;
    int32_t mlp_weight_terms;     /* 0 or 1: weights rounded to bf16 (a network
                                     trained through that rounding);
                                     2: each weight as two bf16 terms, high +
                                     remainder (about 16 bits, an f32-trained
                                     network): one more MFMA per product      */
    int32_t mlp_node_growth;      /* g: spacing growth, 0 .. 1024 (1/256 of the
                                     node's sample index)                     */
    int32_t mlp_node_max;         /* K_max: largest node spacing, K .. 1024
                                     (0: K)                                   */
    int32_t reserved0;            /* 0                                        */
} PstmMigrateArgs;

/* ABI identity, for the loader's layout check. */
int pstm_hip_abi_version(void);
size_t pstm_hip_args_size(void);

/* Message of the last failure on this thread ("" if none). */
const char* pstm_hip_last_error(void);

/* Select the device and create per-device state. Replaces the reference's
 * kernel initialize() (pstm/kernels/base.py:304-305). */
int pstm_hip_init(int device);

/* Release device workspaces. Replaces cleanup() (base.py:317-318). */
int pstm_hip_release(void);

/* Wait for all work queued on `stream`. Replaces synchronize()
 * (base.py:314-315). */
int pstm_hip_stream_sync(void* stream);

/* Launch timing of the gather kernel: while enabled, every gather launch is
 * bracketed by HIP events recorded on its own stream.  reset(1) clears and
 * enables, reset(0) clears and disables; read() waits for the recorded
 * launches and returns their summed duration and count.  Measurement aid
 * with no counterpart in the reference (its benchmark times whole
 * migrate_tile calls on the host, pstm/kernels/factory.py:219-309). */
int pstm_hip_profile_reset(int enable);
int pstm_hip_profile_read(double* total_ms, long long* launches);

/* Launch shape of the gather: mw_cap = sample blocks of 64 per wave
 * (1..4; the workgroup covers 256*mw samples per time window) and batch_cap =
 * traces staged in LDS per batch (1..16); 0 restores the built-in choice.
 * Used by the autotuner; no counterpart in the reference (its tile size
 * heuristic is host-side, pstm/pipeline/tile_planner.py:26-88). */
int pstm_hip_set_tuning(int mw_cap, int batch_cap);

/* Pipelined staging of trace rows: the next batch's rows are fetched while
 * the current batch computes (f32 rows through registers, the batch capped so
 * that its rows fit; anti-aliasing rows by direct global-to-LDS loads into a
 * second row buffer).  Off, each batch's rows go by direct global-to-LDS
 * loads issued at the batch start into one buffer (the batch as large as the
 * LDS allows).  -1 automatic (off), 0 off, 1 on.  Used by the autotuner; no
 * counterpart in the reference. */
int pstm_hip_set_pipeline(int mode);

/* Shape of the most recent gather launch: mw, traces per batch, workgroups
 * launched and work items (pillar groups x time windows). */
int pstm_hip_last_launch(int* mw, int* batch, int* grid, int* items);

/* Migrate one output tile: fold (f64 aperture test, bit-exact with the
 * oracle), image (f32 gather, no atomics) and optionally the effective
 * contribution count E. Replaces migrate_tile (base.py:307-313) and the call
 * of _migrate_tile_kernel_optimized at numba_cpu_optimized.py:593-618. */
int pstm_hip_migrate_tile(const PstmMigrateArgs* args, void* stream);

/* pstm_hip_migrate_tile without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as pstm_hip_migrate_tile, but the call's
 * temporaries (tile-relative geometry, per-sample terms, queue counter,
 * anti-aliasing rows, network operands) live in a fixed 64 MiB device array
 * of the library, one copy per GPU held until the process ends, and the call
 * makes no hipMalloc / hipFree.  Calls whose temporaries exceed 64 MiB are
 * rejected with status 1 (about 32 B per trace, 16 B per (pillar, sample)
 * with 3-D velocity, 8 B per (pillar, sample) with 3-D eta, 8 B per
 * anti-aliasing row sample).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race.  Launch
 * profiling (pstm_hip_profile_reset(1)) must be off. */
int pstm_hip_migrate_tile_noalloc(const PstmMigrateArgs* args, void* stream);

/* Fold only: per-pillar count of traces with midpoint distance <=
 * max_aperture_m (numba_cpu_optimized.py:167-175,231). */
int pstm_hip_fold_tile(const PstmMigrateArgs* args, void* stream);

/* Work-queue migration of several trace groups (common-offset bins) into one
 * shared image in a single persistent launch. groups[g] gives the trace rows
 * [group_start[g], group_start[g+1]) of args->trace_index (or of the block).
 * Replaces the executor's per-bin, per-tile loop
 * (pstm/pipeline/executor.py:977-1358, 1691-1745). */
int pstm_hip_migrate_bins(const PstmMigrateArgs* args, const int32_t* group_start,
                          int32_t n_groups, void* stream);

/* ---------------------------------------------------------------------------
 * Plans: many output tiles (and common-offset bins: a tile's trace list may
 * hold the rows of every bin) in one persistent, hipGraph-captured gather.
 * Replaces the executor's host loop over tiles and bins
 * (pstm/pipeline/executor.py:977-1358 and 1397-1858, dispatch at 1691-1745).
 * ------------------------------------------------------------------------- */
typedef struct PstmTileDesc {
    const double* x_axis;         /* (nx,)                                 */
    const double* y_axis;         /* (ny,)                                 */
    const double* vrms;           /* (nt,) or (nx, ny, nt) per vel_mode    */
    const double* eta;            /* NULL, (nt,) or (nx, ny, nt)           */
    float* image;                 /* (nx, ny, nt) f32, +=                  */
    int32_t* fold;                /* (nx, ny) pillar fold, +=              */
    const int32_t* trace_index;   /* (n_traces,) rows of the block         */
    int32_t n_traces;
    int32_t nx, ny;
    int32_t reserved;             /* 0                                     */
} PstmTileDesc;

typedef struct PstmPlan PstmPlan;

/* Build a plan: `common` holds the trace block (n_traces = its row count),
 * the output time axis and every KernelConfig field; its x_axis, y_axis,
 * vrms, eta, image, fold and trace_index are ignored (per-tile values come
 * from `tiles`).  Geometry, velocity and fold terms are computed here, once;
 * amplitudes and (for the learned traveltime) mlp_params are read at every
 * run.  If common->n_effective is set, every run adds E to it.
 * Each tile must own its image and fold memory: the gather adds into image
 * cells and the fold counts into fold cells without atomics, so the byte
 * ranges of all tiles' images and folds must be pairwise disjoint (bins that
 * share an output tile go into ONE descriptor whose trace_index lists the
 * rows of every bin).  Overlapping ranges are rejected with status 1 before
 * the device is touched. */
int pstm_hip_plan_create(const PstmMigrateArgs* common, const PstmTileDesc* tiles, int32_t n_tiles,
                         PstmPlan** plan);

/* One migration of every tile of the plan on `stream` (image and fold +=).
 * use_graph = 1: the first call captures the run's launch sequence into a
 * hipGraph, later calls replay it; 0: launch directly.  Runs of one plan on
 * one device must not overlap. */
int pstm_hip_plan_run(PstmPlan* plan, int32_t use_graph, void* stream);

/* Work items (tile x pillar group x time window), persistent workgroups and
 * the launch shape of a plan. */
int pstm_hip_plan_info(const PstmPlan* plan, long long* n_items, int* grid, int* mw, int* batch);

int pstm_hip_plan_destroy(PstmPlan* plan);

/* ---------------------------------------------------------------------------
 * Output step: the executor's per-tile header accumulation, gather-bin trace
 * selection and finalize normalisation, on device arrays.  All pointers are
 * device pointers; calls are asynchronous on `stream`.
 * ------------------------------------------------------------------------- */

/* Per-output-bin trace-header sums of one tile, added after the tile as the
 * executor does (pstm/pipeline/executor.py:1783-1806).  For every row r of
 * trace_index (NULL = rows 0 .. n_traces-1) of the block (source_x .. :
 * (rows,) f64), with midpoint m = (s + r) / 2 and offset vector o = r - s:
 *   ix = clip(floor((m_x - x_min) / dx), 0, tile_nx - 1) + x_start,
 *   iy = clip(floor((m_y - y_min) / dy), 0, tile_ny - 1) + y_start,
 *   trace_count[ix, iy] += 1, offset_sum[ix, iy] += |o|,
 *   azimuth_sin_sum[ix, iy] += sin a, azimuth_cos_sum[ix, iy] += cos a,
 *   a = atan2(o_x, o_y) (north = 0; a = 0 for o = 0),
 * into (grid_nx, grid_ny) row-major arrays.  Traces outside the tile (its
 * aperture halo) land in its edge bins, as in the reference.  Sums use
 * atomics: counts are exact, the f64 sums exact up to summation order. */
int pstm_hip_bin_headers(const double* source_x, const double* source_y, const double* receiver_x,
                         const double* receiver_y, const int32_t* trace_index, int32_t n_traces, double x_min,
                         double y_min, double dx, double dy, int32_t x_start, int32_t y_start, int32_t tile_nx,
                         int32_t tile_ny, int32_t grid_nx, int32_t grid_ny, int32_t* trace_count,
                         double* offset_sum, double* azimuth_sin_sum, double* azimuth_cos_sum, void* stream);

/* Gather-bin filter (executor.py:1537-1554): keep[i] = 1 when row
 * trace_index[i] passes every check whose bit is set in `checks`, else 0.
 * offset = |r - s|, o_x = r_x - s_x, o_y = r_y - s_y; bounds are closed, as in
 * the reference (an offset bin is GatherBin.offset_min <= offset <= offset_max,
 * an offset-vector tile bounds o_x and o_y; pstm/config/models.py:677-760). */
#define PSTM_SEL_OFFSET_MIN 1 /* offset >= offset_min */
#define PSTM_SEL_OFFSET_MAX 2 /* offset <= offset_max */
#define PSTM_SEL_OVT_X_MIN 4  /* o_x >= ovt_x_min     */
#define PSTM_SEL_OVT_X_MAX 8  /* o_x <= ovt_x_max     */
#define PSTM_SEL_OVT_Y_MIN 16 /* o_y >= ovt_y_min     */
#define PSTM_SEL_OVT_Y_MAX 32 /* o_y <= ovt_y_max     */
int pstm_hip_select_rows(const double* source_x, const double* source_y, const double* receiver_x,
                         const double* receiver_y, const int32_t* trace_index, int32_t n_traces, int32_t checks,
                         double offset_min, double offset_max, double ovt_x_min, double ovt_x_max, double ovt_y_min,
                         double ovt_y_max, uint8_t* keep, void* stream);

/* Gather-volume normalisation at finalize (executor.py:2155-2160):
 * out[p, k] = image[p, k] / fold[p] where fold[p] > 0, else 0, for pillars
 * p < n_pillars and samples k < nt (the reference's 3-D fold is the pillar
 * fold repeated along t).  f32, correctly rounded; out may equal image. */
int pstm_hip_fold_normalize(const float* image, const int32_t* fold, long long n_pillars, int32_t nt, float* out,
                            void* stream);

/* Time-variant output back onto a uniform axis (resample_to_uniform,
 * pstm/algorithm/time_variant.py:354-397, called at
 * pstm/pipeline/executor.py:1720-1726): for pillars p < n_pillars and output
 * samples k < n_out, out[p, k] = np.interp(out_times_ms[k], tv_times_ms,
 * tv_image[p, :]) computed in f64 with numpy's operations (first / last value
 * outside the time-variant axis) and rounded to f32; tv_image is
 * (n_pillars, n_tv) f32, tv_times_ms must increase.  All zeros when n_tv < 2,
 * as the reference leaves its image then. */
int pstm_hip_resample_uniform(const float* tv_image, const double* tv_times_ms, int32_t n_tv,
                              const double* out_times_ms, int32_t n_out, long long n_pillars, float* out,
                              void* stream);

/* ---------------------------------------------------------------------------
 * Traveltime surrogate (no counterpart in the reference: the learned part of
 * this backend).  A 5-64-64-64-64-2 SiLU MLP; inputs are the encoded
 * (tau, s_s, s_r, v, eta) of PstmMigrateArgs.mlp_*, outputs out_scale * c per
 * leg with t_leg = sqrt(tau^2 + (d/v)^2 (1 + c)).  Parameters: one f32 buffer
 * of PSTM_MLP_PARAMS values, per layer W (out x in, row-major) then b (out).
 * All pointers are device pointers; calls are asynchronous on `stream`.
 * ------------------------------------------------------------------------- */
#define PSTM_MLP_IN 5
#define PSTM_MLP_HIDDEN 64
#define PSTM_MLP_OUT 2
#define PSTM_MLP_PARAMS 12994

int pstm_mlp_param_count(void);
const char* pstm_mlp_last_error(void);

/* Precision of pstm_mlp_forward / _loss_grad / _train_step for the calling
 * process: 0 = f32 (default); 1 = the numerics of the inference fused into the
 * migration (weights rounded to bf16; the input and every hidden activation
 * carried as a bf16 hi+lo pair; f32 accumulation and biases), which makes
 * training quantisation-aware with straight-through gradients. */
int pstm_mlp_set_precision(int bf16);

/* y (n, 2) = network(x (n, 5)), f32. */
int pstm_mlp_forward(const float* params, const float* x, int n, float* y, void* stream);

/* y (n, 2) = network(x (n, 5)) through the device code the migration's
 * learned traveltime uses (BF16 MFMA, f32 accumulation; the numbers of
 * pstm_mlp_forward with pstm_mlp_set_precision(1) up to summation order).
 * Uses the migration library's workspace: do not overlap with a migration
 * call on another stream. */
int pstm_mlp_forward_mfma(const float* params, const float* x, int n, float* y, void* stream);

/* The same with weight_terms = 1 (pstm_mlp_forward_mfma) or 2 (each weight
 * as two bf16 terms, as mlp_weight_terms = 2 in the gather). */
int pstm_mlp_forward_mfma_terms(const float* params, const float* x, int n, float* y, int weight_terms,
                                void* stream);

/* Mean squared error of network(x) against target (n, 2) into *loss (device
 * f32) and its gradient w.r.t. every parameter into grads. */
int pstm_mlp_loss_grad(const float* params, const float* x, const float* target, int n, float* grads,
                       float* loss, void* stream);

/* One Adam step (bias-corrected, step >= 1) on params with moments m, v. */
int pstm_mlp_adam(float* params, const float* grads, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int step, void* stream);

/* loss_grad followed by adam. */
int pstm_mlp_train_step(float* params, float* m, float* v, float* grads, const float* x, const float* target,
                        int n, float lr, float beta1, float beta2, float eps, int step, float* loss,
                        void* stream);

/* The whole training run on `stream`, synchronous: `steps` Adam iterations
 * (bias-corrected step counts step0 + 1 .. step0 + steps) of the mean squared
 * error over minibatches of `batch` rows of (x (n, 5), target (n, 2)), drawn
 * on the device by a counter-based hash of (seed, iteration, sample); the
 * learning rate decays from lr to lr_final on a cosine over the run;
 * iterations >= q_from train through the fused inference's bf16 numerics
 * (quantisation-aware, straight-through gradients).  One fused kernel per
 * iteration computes forward, loss and backward of the minibatch in LDS (no
 * per-layer launches), a second reduces the gradient and applies Adam; runs
 * of iterations are captured once into a hipGraph and replayed.  If loss_log
 * (device f32) is set, it receives the loss of iterations 0, log_every,
 * 2 log_every, ... and of the last one (ceil(steps / log_every) + 1 slots).
 * The training loop the reference would own has no counterpart there (it has
 * no learned traveltime). */
int pstm_mlp_train(float* params, float* m, float* v, const float* x, const float* target, long long n, int steps,
                   int batch, float lr, float lr_final, float beta1, float beta2, float eps, int step0, int q_from,
                   unsigned long long seed, int log_every, float* loss_log, void* stream);

/* Loss and gradient of rows [0, n) by the fused training kernel (the
 * gradient check of pstm_mlp_train), in the precision of
 * pstm_mlp_set_precision. */
int pstm_mlp_loss_grad_fused(const float* params, const float* x, const float* target, int n, float* grads,
                             float* loss, void* stream);

/* Exact one-way VTI times in f64 (pstm/algorithm/anisotropy_vti.py:181-220):
 * t_out[i] = t(tau[i], h[i], v[i], eta[i]); the surrogate's training targets. */
int pstm_vti_leg_times(const double* tau, const double* h, const double* v, const double* eta, long long n,
                       double* t_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PSTM_HIP_H */
