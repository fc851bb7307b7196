// pstm_hip.hip — MI355X (gfx950, CDNA4) PSTM tile migration behind a C ABI.
//
// The reference computes one output tile with a Numba loop nest
// (pstm/kernels/numba_cpu_optimized.py:93-233): pillars in parallel, every
// trace, every output sample, FP64.  Here the same sum is a gather laid out for
// a 64-wide wavefront machine:
//
//   * one workgroup owns NPIL output pillars x a window of output samples and
//     is the only writer of those image cells (no atomics on the image);
//   * lanes of a wave map to consecutive output samples, so the DSR positions
//     of neighbouring lanes land on neighbouring trace samples in LDS
//     (conflict-free reads) while the per-(pillar, trace) terms are uniform;
//   * each wave holds PW pillars x MW sample blocks; the per-(pillar, sample)
//     velocity / aperture / weight terms stay in VGPRs for the whole launch,
//     the image accumulates in VGPRs (f32 within a trace batch, f64 across
//     batches) and is written once;
//   * trace batches are staged HBM -> LDS with coalesced loads, shared by the
//     16 waves of the workgroup; the (pillar, trace) pair table is built once
//     per batch in LDS and a batch no pillar of the workgroup sees is skipped;
//   * the fold is a separate f64 pass so that it is bit-exact with the FP64
//     oracle; absolute coordinates are reduced to tile-relative f32 offsets
//     in f64 before any f32 arithmetic.
//
// Built with -Xclang -target-feature -Xclang -packed-fp32-ops.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pstm_hip.h"

// Shared by the library's translation units (hidden, not exported): the
// last-error slot, the attribute-only and launch-profile switches the gather
// launches read, and the gather launchers of each kernel group.  The groups
// (straight ray with / without anti-aliasing, curved ray, VTI, learned VTI,
// interpolators) are compiled from this file in translation units of their
// own (gather_part1..7.hip, PSTM_GATHER_PART = 1..7) so that the kernel
// variants build in parallel; without PSTM_GATHER_PART this file is the rest
// of the library.
void pstm_set_last_error(const std::string& msg);
bool pstm_gather_attr_only();
bool pstm_gather_prof_on();
void pstm_gather_prof_push(hipEvent_t ev0, hipEvent_t ev1);
int pstm_gather_part1(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);
int pstm_gather_part2(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);
int pstm_gather_part3(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);
int pstm_gather_part4(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);
int pstm_gather_part5(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);
int pstm_gather_part6(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);
int pstm_gather_part7(int mw, bool count, bool lob, int interp, const void* kp, int grid, size_t lds, hipStream_t st);

#ifndef PSTM_GATHER_PART
static __device__ __align__(256) unsigned char scratch[(size_t)64 << 20];
#endif

namespace {

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
#ifndef PSTM_GATHER_PART
thread_local std::string g_last_error;
#endif

int fail(int code, const std::string& msg) {
    pstm_set_last_error(msg);
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return fail(3, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
constexpr int kPW = 2;                 // pillars per wave
constexpr int kTW = 4;                 // time waves per workgroup
constexpr int kPG = 4;                 // pillar waves per workgroup
constexpr int kWaves = kTW * kPG;      // 16 waves
constexpr int kThreads = 64 * kWaves;  // 1024 threads
constexpr int kNPil = kPW * kPG;       // 8 pillars per workgroup
constexpr int kMaxBatch = 16;          // traces per LDS batch (upper bound)
constexpr int kMaxHalfWidth = 64;      // largest anti-aliasing half width (aa_max_half_width)
constexpr int kPF = 12;                // dwords of the next batch's rows each thread holds in flight
constexpr int kStageDw = kMaxBatch * 12;  // staged pair inputs: 12 dwords per trace
constexpr size_t kStageBytes = (size_t)(kStageDw + kMaxBatch) * 4;  // + the block rows after
constexpr size_t kLdsBudget = 160 * 1024 - 2048;  // dynamic LDS; static LDS uses the rest

// traveltimes read from the fused network's t0 lattice
constexpr bool learned_tt(int tt) { return tt == PSTM_TT_VTI_LEARNED || tt == PSTM_TT_CURVED_LEARNED; }

// The t0 lattice of the learned traveltime: nodes n_0 = 0, n_{j+1} = n_j +
// min(max((n_j grow) >> 8, kmin), kmax), the last one clamped to the axis end
// `last` (integer arithmetic, the same on the host, the device and in Python)
__host__ __device__ inline int lattice_next(int n, int kmin, int grow, int kmax, int last) {
    const int step = min(max((n * grow) >> 8, kmin), kmax);
    return min(n + step, last);
}

// per-(pillar, sample) terms, f32, built once per call in f64
//   x : traveltime slowness^2 s = 1/v^2       (curved ray: v(z) at depth)
//   y : weight constant  (spreading ? 1/v : 1) * (obliquity ? t0 : 1)
//   z : taper slope  1/(ap - ts)              (ts = ap (1 - taper))
//   w : taper offset -ts/(ap - ts)  => x_t = d*z + w, in-aperture x_t <= 1
// Per-tile view of the gather's inputs and outputs, in a device table: one
// entry for a single-tile call, one per output tile for a plan
// (pstm_hip_plan_*); the gather looks each work item's tile up.
struct TileView {
    const double* xax;           // (nx) f64 pillar x
    const double* yax;           // (ny) f64 pillar y
    const float4* kc;            // (npv, nt) per-(pillar, sample) terms
    const float2* ek;            // VTI (nev, nt): {2 eta, 1 + 2 eta}
    const double* vrms64;        // (npv, nt) f64 velocity (exact decisions)
    const double* eta64;         // VTI: null (constant), (nt) or (npil, nt)
    float* image;                // (nx, ny, nt) f32, +=
    const float4* geo;           // (ntr, 2): {sx, sy, rx, ry}, {mx, my, w, 0} tile-relative f32
    const int32_t* tidx;         // (ntr) rows of the trace block, or null
    int32_t* fold;               // (nx, ny) i32 (plans: fold += fold_cnt every run)
    const int32_t* fold_cnt;     // plans: (nx, ny) fold counts taken when the plan was built
    int ntr, nx, ny, kc_pstride;
    int ek_pstride, eta_pstride, n_pgroups, item0;  // item0: first work item of the tile
};

// Parameters of one gather launch.
//   kc.x : traveltime slowness^2 s = 1/v^2       (curved ray: v(z) at depth)
//   kc.y : weight constant  (spreading ? 1/v : 1) * (obliquity ? t0 : 1)
//   kc.z : taper slope  1/(ap tf)                => x_t = d * slope + taper_off
struct KParams {
    const TileView* tiles;       // (n_tiles) device table, ascending item0
    int n_tiles;
    unsigned* queue;             // work counter (zeroed before each launch)
    const float* amp;            // (rows, ns)
    const double* prow;          // AA: double-integrated rows
    int prow_global;             // AA rows indexed by block row (plans) or by list position
    const float* qk;             // (nt) tau^2 (curved: v(z)^2)
    const float* t0k;            // (nt) t0 seconds
    unsigned long long* cnt_part;  // (gridDim.x) or null
    int ns, nt;
    int rowlen, lo;              // AA row length / index offset
    int batch;                   // traces per LDS batch
    int pipe;                    // 1: the next batch's rows are staged during the current batch
                                 //    (f32 rows: through registers; AA rows: direct global->LDS
                                 //    loads into the second of two row buffers)
    int rows_bytes;              // bytes of one row buffer
    int pair_off;                // byte offset of the pair table in LDS
    int tot_off;                 // byte offset of the f64 running totals in LDS
    int stage_off;               // byte offset of the staged pair inputs in LDS
    float taper_off;             // x_t = d * slope + taper_off (taper_off = -(1-tf)/tf)
    int n_items;                 // work items over all tiles (pillar groups x time windows)
    int wpow;                    // 0, 1 or 2: power of 1/t in the weight
    int interp;                  // PSTM_INTERP_*
    float idt, uoff, umax;       // u = t*idt + uoff, valid u in [0, umax)
    float umax_lo, umax_hi;      // umax -+ eps_u: band decided in f64
    float max_ap;
    float gdx_idt, gdy_idt, lmax;  // AA
    float ck2, ik_ln2;           // curved: k^2, ln2/k
    float t_floor;               // curved: 0 for k > 0, -inf otherwise
    float v0;                    // curved: surface velocity
    // f64 inputs for the exact decisions near a threshold (see exact_decision)
    const double* sx64;
    const double* sy64;
    const double* rx64;
    const double* ry64;
    const double* mx64;
    const double* my64;
    const double* tax64;         // (nt) ms
    int tt_req;                  // traveltime asked for (PSTM_TT_*)
    float lo_eps;                // eps_u where u >= 0 can be decided differently from f64, else -inf
    float eps_u, eps_x;          // f32 bands around the sample bounds and the aperture edge
    float eps_half;              // relative band around fraction 0.5 (nearest / sinc)
    double tan_dip, min_ap64, max_ap64, t_start_s, dt_s, cv0, ck, eta_const;
    // learned VTI traveltime: network operand images and the t0 lattice
    const void* mlpA;            // bf16 weight operands (k_prep_mlp): high terms, then remainders
    int mlp_split;               // 1: weights as two bf16 terms (mlp_weight_terms = 2)
    const void* mlpB;            // f32 biases in accumulator layout
    int lat_off;                 // byte offset of the lattice in LDS
    int lat_nodes;               // largest lattice node count of a time window
    int lat_cimin_off;           // byte offset of the per-(pillar, node) widest-aperture table
    int lat_list_off;            // byte offset of the list of lattice rows to evaluate
    int node_off;                // byte offset of the per-(pillar, node) network inputs
    int mlp_off;                 // byte offset of the weight operands (then the biases) in LDS
    int lat_kmin, lat_grow, lat_kmax;  // node spacing law (lattice_next)
    int lat_total;               // lattice nodes over the whole output axis
    int lat_pos_off;             // byte offset of the node positions (output samples) in LDS
    int coef_off;                // byte offset of the per-(pair, segment) Hermite coefficients in LDS
    float m_tau_sc, m_v_lo, m_v_sc, m_eta_sc, m_inv_out;
};

__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }

// 0.5 (1 + cos(pi x)) for x in [0, 1] (the reference's cosine taper,
// numba_cpu_optimized.py:211-215) with the hardware cosine, whose argument is
// in revolutions (cos(2 pi r), r = x / 2 in [0, 0.5]): one transcendental
// and two VALU operations; measured against the f64 taper over [0, 1]: at
// most 9.2e-8 absolute
__device__ __forceinline__ float taper_weight(float x) {
    return fmaf(0.5f, __builtin_amdgcn_cosf(0.5f * x), 0.5f);
}

// ---------------------------------------------------------------------------
// preparation kernels (f64)
// ---------------------------------------------------------------------------

#ifndef PSTM_GATHER_PART
// absolute f64 survey coordinates -> tile-relative f32 offsets
__global__ void k_prep_geometry(PstmMigrateArgs a, float4* geo) {
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= a.n_traces) return;
    const int j = a.trace_index ? a.trace_index[jj] : jj;
    const double x0 = a.x_axis[0], y0 = a.y_axis[0];
    const float w = a.trace_weights ? (float)a.trace_weights[j] : 1.0f;
    geo[2 * jj] = make_float4((float)(a.source_x[j] - x0), (float)(a.source_y[j] - y0),
                              (float)(a.receiver_x[j] - x0), (float)(a.receiver_y[j] - y0));
    geo[2 * jj + 1] = make_float4((float)(a.midpoint_x[j] - x0), (float)(a.midpoint_y[j] - y0), w, 0.0f);
}

// per-(pillar|0, sample) terms: _precompute_time_terms
// (numba_cpu_optimized.py:375-401) plus the weight and taper constants.
// tt_eff: traveltime actually evaluated (curved with |k| < 1e-10 becomes
// straight ray in v0, curved_ray.py:35-38).
__global__ void k_prep_time_terms(PstmMigrateArgs a, int tt_eff, int npv, double tan_dip, float4* kc,
                                  float* qk, float* t0k) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = a.nt;
    if (e >= (long)npv * nt) return;
    const int k = (int)(e % nt);
    const double t_s = a.t_axis_ms[k] / 1000.0;
    const double v = a.vrms[e];
    double apv = v * t_s * 0.5 * tan_dip;
    if (apv < a.min_aperture_m) apv = a.min_aperture_m;
    else if (apv > a.max_aperture_m) apv = a.max_aperture_m;
    double s, wv, q;
    if (a.traveltime == PSTM_TT_CURVED || a.traveltime == PSTM_TT_CURVED_LEARNED) {
        const double v0 = a.curved_v0, kg = a.curved_k;
        const double tau = t_s / 2.0;
        const double z = fabs(kg) < 1e-10 ? v0 * tau : (v0 / kg) * (exp(kg * tau) - 1.0);
        const double vz = v0 + kg * z;
        wv = vz;
        if (a.traveltime == PSTM_TT_CURVED_LEARNED) {
            // learned legs: t = z / sqrt(vg^2 + k^2 h^2) (1 + c), vg = sqrt(v0 v(z))
            s = z;
            q = v0 * vz;  // vg^2
        } else if (tt_eff == PSTM_TT_STRAIGHT) {  // k ~ 0: straight ray in v0
            s = 1.0 / (v0 * v0);
            q = tau * tau;
        } else {
            s = vz;        // v(z) at the image depth
            q = vz * vz;   // v(z)^2
        }
    } else {
        s = 1.0 / (v * v);
        wv = 1.0 / sqrt(s);
        q = (t_s * 0.5) * (t_s * 0.5);
    }
    double wc = 1.0;
    if (a.apply_spreading) wc *= 1.0 / wv;
    if (a.apply_obliquity) wc *= t_s;
    // taper coordinate x_t = (d - ap (1 - tf)) / (ap tf) = d / (ap tf) - (1 - tf) / tf:
    // the slope depends on the aperture, the offset only on tf (a kernel scalar)
    const double tf = a.taper_fraction;
    const double slope = tf > 0.0 ? 1.0 / (apv * tf) : 1e30 / apv;
    kc[e] = make_float4((float)s, (float)wc, (float)slope, 0.0f);
    if (qk && e < nt) {  // shared by every tile: written by one tile's preparation
        qk[k] = (float)q;
        t0k[k] = (float)t_s;
    }
}

__global__ void k_prep_eta(PstmMigrateArgs a, int nev, float2* ek) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nev * a.nt) return;
    const double eta = a.eta_mode == PSTM_ETA_CONST ? a.eta_constant : a.eta[e];
    ek[e] = make_float2((float)(2.0 * eta), (float)(1.0 + 2.0 * eta));
}
#endif  // PSTM_GATHER_PART

// f64 distance rounded operation by operation, as the oracle's unfused
// arithmetic: the __d*_rn intrinsics alone do not keep the compiler from
// fusing a product and a sum into one FMA (HIP contracts by default).
__device__ __forceinline__ double dist64(double dx, double dy) {
#pragma clang fp contract(off)
    return __dsqrt_rn(dx * dx + dy * dy);
}

#ifndef PSTM_GATHER_PART
// Fold: count of traces with f64 midpoint distance <= max aperture, one
// workgroup per pillar (numba_cpu_optimized.py:167-175,231).  The products and
// sums are kept unfused so the test is the oracle's, bit for bit.
__global__ __launch_bounds__(256) void k_fold(PstmMigrateArgs a) {
    const int pil = blockIdx.x;
    const int ix = pil / a.ny, iy = pil % a.ny;
    const double ox = a.x_axis[ix], oy = a.y_axis[iy];
    int c = 0;
    for (int jj = threadIdx.x; jj < a.n_traces; jj += blockDim.x) {
        const int j = a.trace_index ? a.trace_index[jj] : jj;
        const double dx = __dsub_rn(ox, a.midpoint_x[j]);
        const double dy = __dsub_rn(oy, a.midpoint_y[j]);
        const double d = dist64(dx, dy);
        c += (d <= a.max_aperture_m) ? 1 : 0;
    }
    __shared__ int red[256];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.fold[pil] += red[0];
}

// Double running sum rows for triangle anti-aliasing, in f64:
//   S[n] = sum_{m<=n} A[m],  P[n] = sum_{m<n} S[m],
//   row[n + lo] = P[n] for n in [-lo, ns + lmax], P = 0 for n <= 0,
//   P[n] = P[ns] + (n - ns) S[ns-1] for n > ns (zero trace outside).
__global__ __launch_bounds__(256) void k_integrate(PstmMigrateArgs a, int rowlen, int lo, double* prow) {
    const int jj = blockIdx.x;
    const int j = a.trace_index ? a.trace_index[jj] : jj;
    const int ns = a.n_samples;
    const float* tr = a.amplitudes + (size_t)j * ns;
    double* out = prow + (size_t)jj * rowlen;
    const int chunk = (ns + 255) / 256;
    const int b = threadIdx.x * chunk, e = min(b + chunk, ns);
    __shared__ double sh[256];
    __shared__ double tot_s, tot_p;
    // pass 1: S
    double ls = 0.0;
    for (int n = b; n < e; ++n) ls += (double)tr[n];
    sh[threadIdx.x] = ls;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {  // inclusive Hillis-Steele scan
        double v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    const double s_before = sh[threadIdx.x] - ls;
    if (threadIdx.x == 255) tot_s = sh[255];
    __syncthreads();
    // pass 2: sum of S over the chunk, scan
    double s = s_before, lp = 0.0;
    for (int n = b; n < e; ++n) {
        s += (double)tr[n];
        lp += s;
    }
    sh[threadIdx.x] = lp;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        double v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0.0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    double p = sh[threadIdx.x] - lp;
    if (threadIdx.x == 255) tot_p = sh[255];
    s = s_before;
    for (int n = b; n < e; ++n) {
        s += (double)tr[n];
        p += s;
        out[n + 1 + lo] = p;  // P[n+1]
    }
    __syncthreads();
    // P[n] for n <= 0
    for (int n = threadIdx.x; n <= lo; n += blockDim.x) out[n] = 0.0;
    // linear continuation beyond the trace
    const double pend = tot_p, send = tot_s;
    for (int n = ns + 1 + (int)threadIdx.x; n + lo < rowlen; n += blockDim.x)
        out[n + lo] = pend + (double)(n - ns) * send;
}

#endif  // PSTM_GATHER_PART

// ---------------------------------------------------------------------------
// exact decisions
// ---------------------------------------------------------------------------
// The reference's sample-bounds test (t_sample < 0 or >= n_samples-1 skips the
// contribution, numba_cpu_optimized.py:199-201) is a discontinuity: an f32
// position that lands on the other side of it than the f64 one drops or adds
// a whole term.  When the f32 position is within eps of a threshold the
// decision is taken again here in f64, with the oracle's formulas and
// operation order, for the aperture test as well (dist_mid > aperture,
// :185-187).  Returns whether the contribution counts and its f64 position.
template <int TT>
__device__ __forceinline__ bool exact_decision(const KParams& P, const TileView& T, int pil, int row, int k,
                                               double* u_out) {
    // every product and sum rounded on its own, as in the oracle
#pragma clang fp contract(off)
    const int ix = pil / T.ny, iy = pil % T.ny;
    const double ox = T.xax[ix], oy = T.yax[iy];
    const double dxm = ox - P.mx64[row], dym = oy - P.my64[row];
    const double dist_mid = dist64(dxm, dym);
    const double v = T.vrms64[(size_t)pil * T.kc_pstride + k];
    const double t_s = P.tax64[k] / 1000.0;
    double ap = v * t_s * 0.5 * P.tan_dip;
    if (ap < P.min_ap64) ap = P.min_ap64;
    else if (ap > P.max_ap64) ap = P.max_ap64;
    const double dxs = ox - P.sx64[row], dys = oy - P.sy64[row];
    const double dxr = ox - P.rx64[row], dyr = oy - P.ry64[row];
    const double ds2 = dxs * dxs + dys * dys;
    const double dr2 = dxr * dxr + dyr * dyr;
    double t;
    if (TT == PSTM_TT_STRAIGHT) {
        // straight ray (numba_cpu_optimized.py:192-194); a curved-ray request
        // with |k| < 1e-10 is the straight ray in v0 (curved_ray.py:35-38)
        const double vt = P.tt_req == PSTM_TT_CURVED ? P.cv0 : v;
        const double sk = 1.0 / (vt * vt);
        const double q = (t_s * 0.5) * (t_s * 0.5);
        t = sqrt(q + ds2 * sk) + sqrt(q + dr2 * sk);
    } else if (TT == PSTM_TT_CURVED) {
        // curved ray (curved_ray.py:19-52, 239-262)
        const double v0 = P.cv0, kg = P.ck, tau = t_s / 2.0;
        const double hs = sqrt(ds2), hr = sqrt(dr2);
        const double z = (v0 / kg) * (exp(kg * tau) - 1.0);
        const double vz = v0 + kg * z;
        const double k1 = kg * hs, k2 = kg * hr;
        t = (1.0 / kg) * log((vz + sqrt(vz * vz + k1 * k1)) / (v0 + sqrt(v0 * v0 + k1 * k1))) +
            (1.0 / kg) * log((vz + sqrt(vz * vz + k2 * k2)) / (v0 + sqrt(v0 * v0 + k2 * k2)));
    } else {
        // VTI one-way legs (anisotropy_vti.py:181-220)
        const double eta = T.eta64 ? T.eta64[(size_t)pil * T.eta_pstride + k] : P.eta_const;
        const double vv = 1.0 / sqrt(1.0 / (v * v));
        const double tau = t_s / 2.0;
        t = 0.0;
        for (int leg = 0; leg < 2; ++leg) {
            const double h = sqrt(leg == 0 ? ds2 : dr2);
            double tl;
            if (vv <= 0 || tau <= 0) {
                tl = tau;
            } else {
                const double h2 = h * h, v2 = vv * vv, t02 = tau * tau;
                const double t2h = t02 + h2 / v2;
                const double den = v2 * (t02 * v2 + (1.0 + 2.0 * eta) * h2);
                tl = fabs(den) < 1e-20 ? sqrt(fmax(t2h, 0.0))
                                       : sqrt(fmax(t2h - (2.0 * eta * h2 * h2) / den, 0.0));
            }
            t += tl;
        }
    }
    const double u = (t - P.t_start_s) / P.dt_s;
    *u_out = u;
    return !(dist_mid > ap) && !(u < 0.0 || u >= (double)(P.ns - 1));
}

// ---------------------------------------------------------------------------
// the gather
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// fused traveltime surrogate: BF16 MFMA, f32 accumulation
// ---------------------------------------------------------------------------
// One wavefront evaluates the network for 16 rows with gfx950's
// v_mfma_f32_16x16x32_bf16 in the transposed form D = W . X^T: A = weights
// (16 outputs x 32 inputs), B = activations (32 inputs x 16 rows), D = 16
// outputs x 16 rows.  Lane l holds row l % 16 and, of an operand, 8 of the
// K step's 32 inputs (lane group g = l / 16: the same 8 in A and B, so the
// products pair up whatever order the hardware sums them in); its
// accumulator register r holds output 4 g + r of the M tile.  The registers
// of M tiles 2 j and 2 j + 1 are therefore taken, as they are, as the B
// operand of K step j of the next layer, and the next layer's weights are
// laid out for that order of its inputs (mlp_unit): activations never leave
// registers.  Sixteen rows (not 32 with the 32 x 32 form) keep the
// accumulators and operands of a layer in 32 registers, so that the gather
// around the evaluation holds its own per-sample terms without spilling.
// The input and every hidden activation enter as a bf16 pair hi + lo (two
// products per K step); the weights as bf16, or as two bf16 terms (one more
// product).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMlpLayers = 5;
__host__ __device__ constexpr int mlp_in(int l) { return l == 0 ? PSTM_MLP_IN : PSTM_MLP_HIDDEN; }
__host__ __device__ constexpr int mlp_out(int l) { return l == kMlpLayers - 1 ? PSTM_MLP_OUT : PSTM_MLP_HIDDEN; }
__host__ __device__ constexpr int mlp_w_off(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += mlp_in(i) * mlp_out(i) + mlp_out(i);
    return off;
}
static_assert(mlp_w_off(kMlpLayers) == PSTM_MLP_PARAMS, "network layout");
static_assert(PSTM_MLP_HIDDEN == 64 && PSTM_MLP_IN <= 8 && PSTM_MLP_OUT <= 4, "operand tiling");

// weight operands, in bf16x8 units: layer 0 [mt 4][lane 64]; layers 1-3
// [mt 4][j 2][lane 64]; layer 4 (outputs padded to 16) [j 2][lane 64]; the
// remainders of split weights follow as a second image of kAVec8 vectors
constexpr int kA1 = 4 * 64, kA4 = kA1 + 3 * 8 * 64, kAVec8 = kA4 + 2 * 64, kAVals = kAVec8 * 8;
// biases in accumulator layout, in f32x4 units: vector v (layers 0-3:
// 4 layer + mt, layer 4: 16) for lane group g at 4 v + g; the accumulator
// layout depends on the lane only through its group
constexpr int kBVec = 4 * 4 + 1, kBVals = kBVec * 4 * 4;

// hidden unit carried by element e of lane group g of K step j (see above)
__host__ __device__ constexpr int mlp_unit(int j, int g, int e) { return 32 * j + 16 * (e >> 2) + 4 * g + (e & 3); }

__host__ __device__ inline unsigned short bf16_bits(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    u += 0x7fffu + ((u >> 16) & 1u);  // nearest, ties to even
    return (unsigned short)(u >> 16);
}

// the evaluation's scale c (mlp_eval16)
constexpr float kSiluC = -1.44269504088896341f;     // -log2(e)
constexpr float kSiluInvC = -0.693147180559945309f;  // 1 / c = -ln(2)

#ifndef PSTM_GATHER_PART
__global__ void k_prep_mlp(const float* __restrict__ P, unsigned short* __restrict__ A, float* __restrict__ B) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kAVals) {
        const int v = e / 8, el = e % 8;
        int l, mt, j, lane;
        if (v < kA1) {
            l = 0; mt = v / 64; j = 0; lane = v % 64;
        } else if (v < kA4) {
            const int r = v - kA1;
            l = 1 + r / (8 * 64); mt = (r / (2 * 64)) % 4; j = (r / 64) % 2; lane = r % 64;
        } else {
            l = 4; mt = 0; j = (v - kA4) / 64; lane = (v - kA4) % 64;
        }
        const int o = mt * 16 + lane % 16, g = lane / 16;
        const int i = l == 0 ? 8 * g + el : mlp_unit(j, g, el);
        const float w = (o < mlp_out(l) && i < mlp_in(l)) ? P[mlp_w_off(l) + o * mlp_in(l) + i] : 0.0f;
        const unsigned short hi = bf16_bits(w);
        A[e] = hi;
        // the weight's remainder as a second bf16 term (split-weight inference)
        A[kAVals + e] = bf16_bits(w - __builtin_bit_cast(float, (unsigned)hi << 16));
    } else if (e < kAVals + kBVals) {
        const int f = e - kAVals, v = f / 16, g = (f / 4) % 4, r = f % 4;
        const int l = v < 16 ? v / 4 : 4, mt = v < 16 ? v % 4 : 0;
        const int o = mt * 16 + 4 * g + r;
        // biases times c = -log2(e) (the scaled evaluation of mlp_eval16)
        B[f] = o < mlp_out(l) ? kSiluC * P[mlp_w_off(l) + mlp_in(l) * mlp_out(l) + o] : 0.0f;
    }
}
#endif  // PSTM_GATHER_PART

// The evaluation carries every pre-activation and activation scaled by
// c = -log2(e): with z' = c z, SiLU(z) = z / (1 + e^-z) gives
// c SiLU(z) = z' / (1 + 2^z') (no multiply before the exponential).  The
// inputs enter as c x and every bias as c b (k_prep_mlp), so that each
// layer's product with its unchanged weights is c times the true one; the
// outputs are divided by c at the end.  bf16 hi + lo splitting is relative,
// so the scale costs no precision, and no weight is re-rounded.  (kSiluC,
// kSiluInvC above.)
__device__ __forceinline__ float silu_c(float zc) { return zc * frcp(1.0f + __builtin_amdgcn_exp2f(zc)); }

// 8 values as a bf16 pair: hi = bf16(a), lo = bf16(a - hi)
__device__ __forceinline__ void split8(const float (&a)[8], bf16x8_t& hi, bf16x8_t& lo) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        hi[i] = (__bf16)a[i];
        lo[i] = (__bf16)(a[i] - (float)hi[i]);
    }
}

// SiLU of the four M tiles, re-packed as the next layer's two B operands
__device__ __forceinline__ void act_pack(const f32x4_t (&acc)[4], bf16x8_t (&h)[2], bf16x8_t (&l)[2]) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        float a[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = silu_c(acc[2 * j + (i >> 2)][i & 3]);
        split8(a, h[j], l[j]);
    }
}

// one K step: W (x_hi + x_lo), plus w_lo x_hi for split weights (the w_lo
// x_lo term, about 2^-16 of the product, is left out)
template <bool SPLIT, typename PA>
__device__ __forceinline__ f32x4_t mfma_w(PA WA, int idx, bf16x8_t bh, bf16x8_t bl, f32x4_t c) {
    const bf16x8_t a = WA[idx];
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, c, 0, 0, 0);
    if (SPLIT) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(WA[kAVec8 + idx], bh, c, 0, 0, 0);
    return c;
}

// Network outputs (y0, y1) of the row held by this lane (valid in lanes
// 0-15; the four lanes l % 16 carry the same row).  Wave-uniform call, all
// lanes active.  SPLIT: each weight carried as two bf16 terms (WA holds the
// high terms, WA + kAVec8 the remainders).  WA, WB: the operand images in
// global memory or (the fused gather) in LDS.
template <bool SPLIT, typename PA, typename PB>
__device__ __forceinline__ void mlp_eval16(PA WA, PB WB, const float (&x)[5], float& y0, float& y1) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    bf16x8_t h[2], l[2];
    {
        // layer 0: the 5 inputs (times c) are elements 0-4 of lane group 0
        const bool g0 = g == 0;
        const float xin[8] = {g0 ? kSiluC * x[0] : 0.0f, g0 ? kSiluC * x[1] : 0.0f, g0 ? kSiluC * x[2] : 0.0f,
                              g0 ? kSiluC * x[3] : 0.0f, g0 ? kSiluC * x[4] : 0.0f, 0.0f, 0.0f, 0.0f};
        bf16x8_t xh, xl;
        split8(xin, xh, xl);
        f32x4_t acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) acc[mt] = mfma_w<SPLIT>(WA, mt * 64 + lane, xh, xl, WB[mt * 4 + g]);
        act_pack(acc, h, l);
    }
#pragma unroll 1
    for (int layer = 1; layer < 4; ++layer) {
        const int wa = kA1 + (layer - 1) * 8 * 64;
        f32x4_t acc[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            acc[mt] = WB[(layer * 4 + mt) * 4 + g];
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[mt] = mfma_w<SPLIT>(WA, wa + (mt * 2 + j) * 64 + lane, h[j], l[j], acc[mt]);
        }
        act_pack(acc, h, l);
    }
    f32x4_t o = WB[16 * 4 + g];
#pragma unroll
    for (int j = 0; j < 2; ++j) o = mfma_w<SPLIT>(WA, kA4 + j * 64 + lane, h[j], l[j], o);
    y0 = o[0] * kSiluInvC;
    y1 = o[1] * kSiluInvC;
}

#ifndef PSTM_GATHER_PART
// Batched inference through the fused path's device code: one wavefront per
// 16 rows of x (n, 5), outputs y (n, 2).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_mlp_forward_mfma(const bf16x8_t* __restrict__ WA,
                                                          const f32x4_t* __restrict__ WB,
                                                          const float* __restrict__ x, int n, float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const long wv = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long r = wv * 16 + (lane & 15);
    float xi[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (r < n) {
#pragma unroll
        for (int i = 0; i < 5; ++i) xi[i] = x[r * 5 + i];
    }
    float y0, y1;
    mlp_eval16<SPLIT>(WA, WB, xi, y0, y1);
    if (lane < 16 && r < n) {
        y[2 * r] = y0;
        y[2 * r + 1] = y1;
    }
}
#endif  // PSTM_GATHER_PART

// Derivative at node i (1 or 2) of values f[0..3] at node positions x[0..3]
// (ok: the node's value is there): the three-point difference over i - 1,
// i, i + 1 where both neighbours are there, else the one-sided three-point
// difference over i, i -+ 1, i -+ 2, else the two-point one (second order:
// exact for a quadratic).
// (the quotients through one hardware reciprocal each: within ~2 ulp of
// the divided forms, far inside the network's own accuracy)
__device__ __forceinline__ float d3(float a, float b, float c, float xa, float xb, float xc) {
    // derivative at xa of the parabola through (xa, a), (xb, b), (xc, c):
    // [-(2 h1 + h2) h2 a + (h1 + h2)^2 b - h1^2 c] / (h1 h2 (h1 + h2))
    const float h1 = xb - xa, h2 = xc - xb, s = h1 + h2;
    return (-(2.0f * h1 + h2) * h2 * a + s * s * b - h1 * h1 * c) * frcp(h1 * h2 * s);
}

__device__ __forceinline__ float node_derivative(const float (&f)[4], const bool (&ok)[4], const float (&x)[4], int i) {
    if (i == 1) {
        if (ok[0]) {
            const float h1 = x[1] - x[0], h2 = x[2] - x[1];
            return (h1 * h1 * (f[2] - f[1]) + h2 * h2 * (f[1] - f[0])) * frcp(h1 * h2 * (h1 + h2));
        }
        return ok[3] ? d3(f[1], f[2], f[3], x[1], x[2], x[3]) : (f[2] - f[1]) * frcp(x[2] - x[1]);
    }
    if (ok[3]) {
        const float h1 = x[2] - x[1], h2 = x[3] - x[2];
        return (h1 * h1 * (f[3] - f[2]) + h2 * h2 * (f[2] - f[1])) * frcp(h1 * h2 * (h1 + h2));
    }
    return ok[0] ? d3(f[2], f[1], f[0], x[2], x[1], x[0]) : (f[2] - f[1]) * frcp(x[2] - x[1]);
}

// Traveltime of one contribution from the pair terms e/f (LDS) and the
// (pillar, sample) terms.  y1/y2 return 1/t_src, 1/t_rec for anti-aliasing.
// t0z: the sample's t0 may be <= 0 (the VTI times are then t0).
template <int TT, bool AA>
__device__ __forceinline__ float contrib_time(const KParams& P, const float4& e, const float4& f, float q, float cs,
                                              float e2v, float e1v, float t0v, float etav, float& y1, float& y2,
                                              bool t0z = true) {
    float t;
    y1 = 0.f;
    y2 = 0.f;
    if (TT == PSTM_TT_STRAIGHT) {
        const float a1 = fmaf(e.x, cs, q);
        const float a2 = fmaf(e.y, cs, q);
        if (AA) {
            // a >= 0; at a = 0 the leg is 0 (a * y = 0) and so is its dip term
            // (the offsets multiplying y are 0 there)
            y1 = frsq(fmaxf(a1, 1e-30f));
            y2 = frsq(fmaxf(a2, 1e-30f));
            t = a1 * y1 + a2 * y2;
        } else {
            t = fsqrt(a1) + fsqrt(a2);
        }
    } else if (TT == PSTM_TT_CURVED) {
        // t = ln[(vz + sqrt(vz^2 + k^2 hs^2)) (vz + sqrt(vz^2 + k^2 hr^2)) / (den_s den_r)] / k
        // (curved_ray.py:43-52, both legs in one logarithm)
        const float ns_ = cs + fsqrt(q + e.x);
        const float nr_ = cs + fsqrt(q + e.y);
        t = flog2((ns_ * nr_) * f.x) * P.ik_ln2;
        // k > 0: v(z) >= v0, so the exact time is >= 0 (and exactly 0 at t0 = 0,
        // where the f32 ratio only rounds to 1 +- ulp)
        t = fmaxf(t, P.t_floor);
    } else if (TT == PSTM_TT_CURVED_LEARNED) {
        // learned curved legs: t = z / sqrt(vg^2 + k^2 h^2) (1 + c) with the
        // depth z (cs), vg^2 (q), k^2 h^2 (e.x, e.y) and c from the lattice
        // (e2v, e1v); 1 + c >= 0 keeps t >= 0, and t = 0 at t0 = 0 (z = 0)
        t = cs * fmaf(frsq(e.x + q), fmaxf(1.0f + e2v, 0.0f), frsq(e.y + q) * fmaxf(1.0f + e1v, 0.0f));
    } else if (TT == PSTM_TT_VTI_LEARNED) {
        // learned legs: t^2 = tau^2 + h^2 s (1 + eta g), g = c / eta from the
        // lattice (e2v, e1v here), s = cs and etav = s eta at the sample:
        // h^2 (s + s eta g); s (1 + eta g) >= 0 keeps t >= t0 as the exact VTI
        // time is
        const float a1 = fmaf(e.x, fmaxf(fmaf(etav, e2v, cs), 0.0f), q);
        const float a2 = fmaf(e.y, fmaxf(fmaf(etav, e1v, cs), 0.0f), q);
        t = fsqrt(a1) + fsqrt(a2);
        if (t0z && !(t0v > 0.f)) t = t0v;
    } else {
        // VTI one-way legs (anisotropy_vti.py:181-220) with s = 1/v^2:
        // t^2 = tau^2 + h^2 s - 2 eta (h^2 s)^2 / (tau^2 + (1 + 2 eta) h^2 s)
        const float hs = e.x * cs, hr = e.y * cs;
        const float ts2 = q + hs - e2v * hs * hs * frcp(fmaf(e1v, hs, q));
        const float tr2 = q + hr - e2v * hr * hr * frcp(fmaf(e1v, hr, q));
        t = fsqrt(fmaxf(ts2, 0.f)) + fsqrt(fmaxf(tr2, 0.f));
        if (t0z && !(t0v > 0.f)) t = t0v;
    }
    return t;
}

// ---------------------------------------------------------------------------
// trace interpolation other than linear (interpolate_sample,
// interpolation.py:276-318) at sample position u = i0 + f0 of an f32 row in
// LDS, 0 <= i0 <= ns - 2, 0 <= f0 < 1 (both call sites guarantee it).  Every
// threshold of the reference is a test on i0 alone (u < H <=> i0 < H,
// u >= ns - H <=> i0 >= ns - H, int(u + 0.5) = i0 + (f0 >= 0.5)), so each
// index below stays inside the row.  The wider kernels fall back to linear
// near the trace ends, as the reference's do.  Windowed sinc and Lanczos
// need sin(pi x) for the taps x = fr - k: sin(pi (fr - k)) = (-1)^k sin(pi fr),
// and the window / second factor by angle addition with cos, sin(pi k / H)
// from a table, so a contribution costs two transcendental evaluations, not
// two per tap.
// ---------------------------------------------------------------------------
constexpr float kPiF = 3.14159265358979f;

__device__ __forceinline__ float lerp_at(const float* row, int i0, float f0) {
    return row[i0] * (1.0f - f0) + row[i0 + 1] * f0;
}

// {cos(pi k / H), sin(pi k / H)} for k = j - H + 1, j < 2 H
template <int H>
__device__ __forceinline__ float2 tap_trig(int j) {
    if (H == 2) {
        constexpr float c[4] = {0.0f, 1.0f, 0.0f, -1.0f};
        constexpr float s[4] = {-1.0f, 0.0f, 1.0f, 0.0f};
        return make_float2(c[j], s[j]);
    } else if (H == 3) {
        constexpr float c[6] = {-0.5f, 0.5f, 1.0f, 0.5f, -0.5f, -1.0f};
        constexpr float s[6] = {-0.8660253882408142f, -0.8660253882408142f, 0.0f, 0.8660253882408142f,
                                0.8660253882408142f, 0.0f};
        return make_float2(c[j], s[j]);
    } else if (H == 4) {
        constexpr float c[8] = {-0.7071067690849304f, 0.0f, 0.7071067690849304f, 1.0f, 0.7071067690849304f, 0.0f,
                                -0.7071067690849304f, -1.0f};
        constexpr float s[8] = {-0.7071067690849304f, -1.0f, -0.7071067690849304f, 0.0f, 0.7071067690849304f, 1.0f,
                                0.7071067690849304f, 0.0f};
        return make_float2(c[j], s[j]);
    } else if (H == 5) {
        constexpr float c[10] = {-0.80901700258255f, -0.30901700258255005f, 0.30901700258255005f, 0.80901700258255f,
                                 1.0f, 0.80901700258255f, 0.30901700258255005f, -0.30901700258255005f,
                                 -0.80901700258255f, -1.0f};
        constexpr float s[10] = {-0.5877852439880371f, -0.9510565400123596f, -0.9510565400123596f,
                                 -0.5877852439880371f, 0.0f, 0.5877852439880371f, 0.9510565400123596f,
                                 0.9510565400123596f, 0.5877852439880371f, 0.0f};
        return make_float2(c[j], s[j]);
    } else {
        constexpr float c[16] = {-0.9238795042037964f, -0.7071067690849304f, -0.3826834261417389f, 0.0f,
                                 0.3826834261417389f, 0.7071067690849304f, 0.9238795042037964f, 1.0f,
                                 0.9238795042037964f, 0.7071067690849304f, 0.3826834261417389f, 0.0f,
                                 -0.3826834261417389f, -0.7071067690849304f, -0.9238795042037964f, -1.0f};
        constexpr float s[16] = {-0.3826834261417389f, -0.7071067690849304f, -0.9238795042037964f, -1.0f,
                                 -0.9238795042037964f, -0.7071067690849304f, -0.3826834261417389f, 0.0f,
                                 0.3826834261417389f, 0.7071067690849304f, 0.9238795042037964f, 1.0f,
                                 0.9238795042037964f, 0.7071067690849304f, 0.3826834261417389f, 0.0f};
        return make_float2(c[j], s[j]);
    }
}

// Hann-windowed sinc of half width H (_sinc_interp, interpolation.py:154-191):
// taps k = -H+1 .. H around the nearest sample, weights sinc(x) *
// 0.5 (1 + cos(pi x / H)) (0 for |x| >= H), taps past the trace end left out,
// normalised by the sum of the weights
template <int H>
__device__ __forceinline__ float sinc_at(const float* row, int i0, float f0, int ns) {
    if (i0 < H || i0 >= ns - H) return lerp_at(row, i0, f0);
    const bool up = f0 >= 0.5f;
    const int ic = i0 + (up ? 1 : 0);           // in [H, ns - H]
    const float fr = up ? f0 - 1.0f : f0;       // [-0.5, 0.5), exact
    const float sp = sinpif(fr);
    float sa, ca;
    sincospif(fr / (float)H, &sa, &ca);
    float res = 0.0f, norm = 0.0f;
#pragma unroll
    for (int j = 0; j < 2 * H; ++j) {
        const int k = j - H + 1;
        const int i = ic + k;                   // in [1, ns]
        const float x = fr - (float)k;
        const float2 cs = tap_trig<H>(j);
        const float sinc = fabsf(x) < 1e-10f ? 1.0f : ((k & 1) ? -sp : sp) * __builtin_amdgcn_rcpf(kPiF * x);
        const float win = fabsf(x) >= (float)H ? 0.0f : 0.5f * (1.0f + fmaf(ca, cs.x, sa * cs.y));
        const float w = i < ns ? sinc * win : 0.0f;
        res = fmaf(row[min(i, ns - 1)], w, res);
        norm += w;
    }
    return norm > 1e-10f ? res / norm : 0.0f;
}

// Lanczos of order A (_lanczos_interp, interpolation.py:222-256): taps
// k = -A+1 .. A around floor(u), L(x) = sinc(x) sinc(x / A) (1 at x = 0, 0 for
// |x| >= A), normalised by the sum of the weights
template <int A>
__device__ __forceinline__ float lanczos_at(const float* row, int i0, float f0, int ns) {
    if (i0 < A || i0 >= ns - A) return lerp_at(row, i0, f0);
    const float sp = sinpif(f0);
    float sa, ca;
    sincospif(f0 / (float)A, &sa, &ca);
    float res = 0.0f, norm = 0.0f;
#pragma unroll
    for (int j = 0; j < 2 * A; ++j) {
        const int k = j - A + 1;                // i0 + k in [1, ns - 1]
        const float x = f0 - (float)k;
        const float2 cs = tap_trig<A>(j);
        const float snx = (k & 1) ? -sp : sp;               // sin(pi x)
        const float sna = fmaf(sa, cs.x, -ca * cs.y);       // sin(pi x / A)
        const float px = kPiF * x;
        float w = (float)A * snx * sna * __builtin_amdgcn_rcpf(px * px);
        w = fabsf(x) < 1e-10f ? 1.0f : (fabsf(x) >= (float)A ? 0.0f : w);
        res = fmaf(row[i0 + k], w, res);
        norm += w;
    }
    return norm > 1e-10f ? res / norm : 0.0f;
}

// sample blocks per wave of each interpolator's one instantiation (16 sinc
// taps leave room for one block of accumulators)
template <int INTERP>
constexpr int kInterpMW = INTERP == PSTM_INTERP_SINC16 ? 1 : 2;
inline int interp_mw(int interp) { return interp == PSTM_INTERP_SINC16 ? 1 : 2; }

// interpolators whose value jumps where the fraction crosses 0.5
template <int INTERP>
constexpr bool kHalfJump = INTERP == PSTM_INTERP_NEAREST || INTERP == PSTM_INTERP_SINC4 ||
                           INTERP == PSTM_INTERP_SINC8 || INTERP == PSTM_INTERP_SINC16;

template <int INTERP>
__device__ __forceinline__ float interp_at(const float* row, int i0, float f0, int ns) {
    if (INTERP == PSTM_INTERP_NEAREST) {
        return row[i0 + (f0 >= 0.5f ? 1 : 0)];  // int(u + 0.5) (_nearest_interp, :58-65)
    } else if (INTERP == PSTM_INTERP_CUBIC) {
        // Catmull-Rom on 4 samples (_cubic_interp, interpolation.py:95-135)
        if (i0 < 1 || i0 >= ns - 2) return lerp_at(row, i0, f0);
        const float p0 = row[i0 - 1], p1 = row[i0], p2 = row[i0 + 1], p3 = row[i0 + 2];
        const float a0 = -0.5f * p0 + 1.5f * p1 - 1.5f * p2 + 0.5f * p3;
        const float a1 = p0 - 2.5f * p1 + 2.0f * p2 - 0.5f * p3;
        const float a2 = -0.5f * p0 + 0.5f * p2;
        return fmaf(fmaf(fmaf(a0, f0, a1), f0, a2), f0, p1);
    } else if (INTERP == PSTM_INTERP_SINC4) {
        return sinc_at<2>(row, i0, f0, ns);
    } else if (INTERP == PSTM_INTERP_SINC8) {
        return sinc_at<4>(row, i0, f0, ns);
    } else if (INTERP == PSTM_INTERP_SINC16) {
        return sinc_at<8>(row, i0, f0, ns);
    } else if (INTERP == PSTM_INTERP_LANCZOS3) {
        return lanczos_at<3>(row, i0, f0, ns);
    } else {
        return lanczos_at<5>(row, i0, f0, ns);
    }
}

// some lane of the wave holds p: the ballot of the predicate itself, so the
// mask stays in scalar registers (__any takes an int, which costs a select and
// a compare per call)
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }
__device__ __forceinline__ unsigned long long wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// Workgroup barrier for data in LDS: this wave's LDS accesses complete
// (lgkmcnt(0)), then s_barrier, and no memory access moves across it.  Unlike
// __syncthreads it does not wait for vector memory operations: the fence of
// __syncthreads counts direct global -> LDS loads as LDS writes and waits
// for them all (vmcnt(0)), so row loads issued early would land before the
// next barrier of every phase.  A direct load's own wave waits for it
// (vmcnt(0)) before a barrier after which other waves read its data.
// a / d for 0 <= a < 2^16 and 2 <= d < 2^16 by one multiply-high with
// m = div_magic(d) = floor((2^32 - 1) / d) + 1: a m / 2^32 exceeds a / d by
// less than a / 2^32 < 2^-16, less than the distance 1 / d from a / d's
// fraction to the next integer, so the floor is exact.  The magic number of
// a wave-uniform divisor is kept in a scalar register (a divisor held in a
// vector register across the batch loop was spilled to scratch memory, whose
// reload waits for every vector memory load in flight).
__device__ __forceinline__ unsigned div_magic(int d) {
    return __builtin_amdgcn_readfirstlane(d > 1 ? 0xFFFFFFFFu / (unsigned)d + 1u : 0u);
}
__device__ __forceinline__ int div_by(int a, unsigned m, int d) { return d > 1 ? (int)__umulhi((unsigned)a, m) : a; }
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Trace amplitude at sample position i0 + f0: linear interpolation
// (interpolation.py:68-92) or another interpolator (interp_at), or, with
// anti-aliasing, the triangle filter of half width L (from the operator time
// dip) on the f64 double running sum: a blend of the integer half widths
// floor(L) and floor(L)+1, each a second difference interpolated between i0
// and i0+1.
//
// The amplitude is split into a fetch (every LDS read it needs) and a finish
// (the arithmetic on what was read), so that the gather can issue the reads
// first and work out the weight while they are in flight.
struct LinTaps {
    float s0, s1;
};
struct AaTaps {
    double am1, am0, am1p, c0, c1, b0, b1, b2;  // running sum at i0 - lf - 1 .. i0 + lf + 2
    float2 iw;                                  // {1 / lf^2, 1 / (lf + 1)^2}
    float lam;                                  // fraction of L between lf and lf + 1
};

__device__ __forceinline__ LinTaps lin_fetch(const float* row, int i0) {
    const float* pc = row + i0;
    return LinTaps{pc[0], pc[1]};
}

__device__ __forceinline__ float lin_finish(const LinTaps& s, float f0) { return fmaf(f0, s.s1 - s.s0, s.s0); }

// half width L = operator time dip x bin size in samples (offsets in f carry
// the bin size / dt), clamped to [1, lmax]; the taps of floor(L) and
// floor(L) + 1 around i0 and i0 + 1, addressed from one centre pointer
__device__ __forceinline__ AaTaps aa_fetch(const KParams& P, const double* row, int i0, const float4& f, float cs,
                                           float y1, float y2, const float2* invsq) {
    const float gx = fmaf(f.x, y1, f.z * y2);
    const float gy = fmaf(f.y, y1, f.w * y2);
    const float L = __builtin_amdgcn_fmed3f(cs * fmaxf(fabsf(gx), fabsf(gy)), 1.0f, P.lmax);
    const int lf = (int)L;
    AaTaps a;
    a.lam = __builtin_amdgcn_fractf(L);
    // pc points one element before the centre, so that every tap is at a
    // non-negative offset of pc, pc - lf or pc + lf (LDS instruction offsets
    // are unsigned); the address is hidden from the optimiser, which would
    // otherwise expand pc - lf and pc + lf from i0 and lf separately (six
    // integer operations instead of four)
    using LdsDouble = __attribute__((address_space(3))) const double;
    LdsDouble* pc = (LdsDouble*)(row + P.lo - 1) + i0;
    asm("" : "+v"(pc));
    LdsDouble* pa = pc - lf;
    LdsDouble* pb = pc + lf;
    a.am1 = pa[0];
    a.am0 = pa[1];
    a.am1p = pa[2];
    a.c0 = pc[1];
    a.c1 = pc[2];
    a.b0 = pb[1];
    a.b1 = pb[2];
    a.b2 = pb[3];
    a.iw = invsq[lf];
    return a;
}

// triangle filters of half widths lf and lf + 1 at i0 and i0 + 1, blended by
// lam across L and by f0 across the sample
__device__ __forceinline__ float aa_finish(const AaTaps& a, float f0) {
    // second differences P[n + L] - 2 P[n] + P[n - L] (one add, one fma in f64)
    const float dl0 = (float)fma(-2.0, a.c0, a.b0 + a.am0);   // T_lf[i0] * lf^2
    const float dh0 = (float)fma(-2.0, a.c0, a.b1 + a.am1);   // T_lf+1[i0] * (lf+1)^2
    const float dl1 = (float)fma(-2.0, a.c1, a.b1 + a.am1p);  // T_lf[i0+1] * lf^2
    const float dh1 = (float)fma(-2.0, a.c1, a.b2 + a.am0);   // T_lf+1[i0+1] * (lf+1)^2
    const float wl = fmaf(-a.lam, a.iw.x, a.iw.x);
    const float wh = a.lam * a.iw.y;
    const float tr0 = fmaf(dl0, wl, dh0 * wh);
    const float tr1 = fmaf(dl1, wl, dh1 * wh);
    return fmaf(f0, tr1 - tr0, tr0);
}

template <bool AA, int INTERP, typename Row>
__device__ __forceinline__ float contrib_amp(const KParams& P, const Row* row, int i0, float f0, const float4& f,
                                             float cs, float y1, float y2, const float2* invsq) {
    if constexpr (AA) {
        return aa_finish(aa_fetch(P, row, i0, f, cs, y1, y2, invsq), f0);
    } else if constexpr (INTERP != PSTM_INTERP_LINEAR) {
        return interp_at<INTERP>(row, i0, f0, P.ns);
    } else {
        return lin_finish(lin_fetch(row, i0), f0);
    }
}

// Weight: trace weight x cosine taper x spreading 1/(v t) x obliquity t0/t,
// the last two only for t > 1 ms (numba_cpu_optimized.py:209-226).
__device__ __forceinline__ float contrib_weight(const KParams& P, float wtr, float xt, float t, float cw) {
    float w = wtr;
    if (xt > 0.0f) w *= taper_weight(fminf(xt, 1.0f));
    if (P.wpow != 0 && t > 0.001f) {
        const float r = frcp(t);
        w *= cw * (P.wpow == 2 ? r * r : r);
    }
    return w;
}

// W2: spreading and obliquity both on (weight factor cw / t^2, the common case)
// as a compile-time constant; otherwise the power of 1/t (0 or 1) is read from P
template <int MW, int TT, bool AA, bool COUNT, bool LOB, bool W2, int INTERP = PSTM_INTERP_LINEAR>
__global__ __launch_bounds__(kThreads) void k_migrate(KParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    using Row = typename std::conditional<AA, double, float>::type;
    constexpr int kPairVec = (AA || TT == PSTM_TT_CURVED) ? 2 : 1;
    Row* rows = reinterpret_cast<Row*>(smem);  // current row buffer
    float4* pairs = reinterpret_cast<float4*>(smem + P.pair_off);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // the wave index is wave-uniform; say so, so that everything derived from
    // it (pair indices, the pair data read from LDS, their branches) is scalar
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tw = wave % kTW;
    const int pg = wave / kTW;
    const int win = 64 * kTW * MW;
    const bool pipe = !learned_tt(TT) && P.pipe;
    unsigned long long cnt = 0;
    __shared__ int s_item;
    // block rows of the next batch (its row loads and pair table read them
    // from here), in two copies by batch parity: a
    // skipped batch has one barrier only, after which the waves read the copy
    // of the batch while a wave that runs ahead writes the other one for the
    // batch after; the copy is written again only after the next barrier
    __shared__ int s_rownext[2][kMaxBatch];
    __shared__ float2 s_invsq[kMaxHalfWidth + 2];  // AA: {1 / l^2, 1 / (l + 1)^2}
    __shared__ int s_nrows;  // learned traveltime: lattice rows listed for evaluation
    // per batch parity: the batch's first trace j0 when some pair of the batch
    // is inside the maximum aperture (written by the waves holding such a pair)
    __shared__ int s_any[2];
    // learned traveltime: per pillar of the group its Lagrange weights on
    // pillars 0, p_mid and p_hi (the item's, written at the item start)
    __shared__ float4 s_lw[kNPil];
    // the next batch's pair inputs, loaded during the batch before by direct
    // global -> LDS loads of wave 0, which builds the pair table: per trace 12
    // dwords, the geometry {sx, sy, rx, ry}, {mx, my, w, 0}
    // and the f64 midpoint (x, y); and the block rows of the batch after it
    // (in the dynamic LDS, after the pair table: arrays written by direct
    // loads alone must not be static ones, whose reads the compiler may fold
    // as never written)
    uint32_t* s_stage = reinterpret_cast<uint32_t*>(smem + P.stage_off);
    int* s_stage_row = reinterpret_cast<int*>(smem + P.stage_off + kStageDw * 4);
    if (AA) {
        for (int i = tid; i < kMaxHalfWidth + 2; i += kThreads)
            s_invsq[i] = make_float2(i > 0 ? 1.0f / (float)(i * i) : 0.0f, 1.0f / (float)((i + 1) * (i + 1)));
    }  // visible after the first barrier of the item loop
    // learned traveltime: the network's operand images stay in LDS for the
    // workgroup's life (a weight operand per MFMA from global memory would
    // wait out the L2 latency once per product)
    using LdsA = __attribute__((address_space(3))) const bf16x8_t;
    using LdsB = __attribute__((address_space(3))) const f32x4_t;
    const int mlp_nv = learned_tt(TT) ? kAVec8 * (P.mlp_split ? 2 : 1) : 0;
    if (learned_tt(TT)) {
        if (tid == 0) {
            int* npos = reinterpret_cast<int*>(smem + P.lat_pos_off);
            npos[0] = 0;
            for (int j = 1; j < P.lat_total; ++j)
                npos[j] = lattice_next(npos[j - 1], P.lat_kmin, P.lat_grow, P.lat_kmax, P.nt - 1);
        }
        const uint4* ws = static_cast<const uint4*>(P.mlpA);
        uint4* wd = reinterpret_cast<uint4*>(smem + P.mlp_off);
        for (int i = tid; i < mlp_nv; i += kThreads) wd[i] = ws[i];
        const float4* bs = static_cast<const float4*>(P.mlpB);
        float4* bd = reinterpret_cast<float4*>(smem + P.mlp_off + mlp_nv * 16);
        for (int i = tid; i < kBVals / 4; i += kThreads) bd[i] = bs[i];
    }  // visible after the first barrier of the item loop

    // persistent workgroups drain the work queue; every workgroup leaves when
    // the counter passes n_items
    for (;;) {
        if (tid == 0) s_item = (int)atomicAdd(P.queue, 1u);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);  // uniform: keeps the tile view scalar
        if (item >= P.n_items) break;
        int lo_t = 0, hi_t = P.n_tiles - 1;  // last tile with item0 <= item
        while (lo_t < hi_t) {
            const int mid = (lo_t + hi_t + 1) >> 1;
            if (P.tiles[mid].item0 <= item) lo_t = mid;
            else hi_t = mid - 1;
        }
        // read in place (scalar loads where used) rather than copied to registers
        const TileView& T = P.tiles[lo_t];
        const int npil = T.nx * T.ny;
        const int li = item - T.item0;
        const int pgroup = li % T.n_pgroups;
        const int kwin = li / T.n_pgroups;
        const int pil0 = pgroup * kNPil;
        const int kbase = kwin * win;
        // learned traveltime: the network is evaluated at three pillars of the
        // group, 0, p_mid and p_hi (the last valid one)
        const int p_hi = min(kNPil - 1, npil - 1 - pil0);
        const int p_mid = (p_hi + 1) >> 1;

        // register-resident per-(pillar, sample) terms
        float q[MW], cs[kPW][MW], ci[kPW][MW];
        float e2[kPW][MW], e1[kPW][MW], t0r[MW];
        float etar[kPW][MW];  // learned VTI: slowness^2 x eta at (pillar, sample)
        float acc[kPW][MW];
        // running totals across batches: f64 lane-private in LDS for the
        // exact traveltimes, f32 in registers for the learned one (whose
        // accuracy is that of the network; the LDS goes to larger batches)
        double* tot = reinterpret_cast<double*>(smem + P.tot_off);  // [kPW*MW][kThreads]
        float totf[kPW][MW];
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            const int k = kbase + 64 * (tw + kTW * m) + lane;
            const bool kv = k < P.nt;
            q[m] = kv ? P.qk[k] : 0.0f;
            t0r[m] = kv ? P.t0k[k] : 0.0f;
#pragma unroll
            for (int p = 0; p < kPW; ++p) {
                const int pil = pil0 + pg * kPW + p;
                acc[p][m] = 0.0f;
                if (learned_tt(TT)) totf[p][m] = 0.0f;
                else tot[(p * MW + m) * kThreads + tid] = 0.0;
                e2[p][m] = 0.0f;
                e1[p][m] = 1.0f;
                etar[p][m] = 0.0f;
                if (kv && pil < npil) {
                    const float4 c = T.kc[(size_t)pil * T.kc_pstride + k];
                    cs[p][m] = c.x;
                    ci[p][m] = c.z;
                    if (TT == PSTM_TT_VTI) {
                        const float2 ee = T.ek[(size_t)pil * T.ek_pstride + k];
                        e2[p][m] = ee.x;
                        e1[p][m] = ee.y;
                    } else if (TT == PSTM_TT_VTI_LEARNED) {
                        etar[p][m] = c.x * (0.5f * T.ek[(size_t)pil * T.ek_pstride + k].x);  // s eta
                    }
                } else {
                    cs[p][m] = 0.0f;
                    ci[p][m] = INFINITY;  // x_t = +inf (d > 0) or NaN (d = 0): never in aperture
                }
            }
        }

        // the weight constant kc.y of a slot: constant over the traces, so
        // applied to the slot's total when the item ends; only early blocks
        // (below) need it per contribution, where it is read when used
        auto cw_at = [&](int p, int m) -> float {
            const int k = kbase + 64 * (tw + kTW * m) + lane;
            const int pil = pil0 + pg * kPW + p;
            return (k < P.nt && pil < npil) ? T.kc[(size_t)pil * T.kc_pstride + k].y : 0.0f;
        };

        // sample blocks of this wave whose first t0 is <= 2 ms: only there can
        // t <= 1 ms switch the spreading/obliquity weights off (wave-uniform)
        unsigned early = 0;
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            const int kb = kbase + 64 * (tw + kTW * m);
            if (kb < P.nt && P.t0k[kb] <= 0.002f) early |= 1u << m;
        }
        if (TT == PSTM_TT_CURVED || TT == PSTM_TT_CURVED_LEARNED) early = (1u << MW) - 1u;  // t may be < t0 on a curved ray
        early = __builtin_amdgcn_readfirstlane(early);

        // the pair table, pair pt = p * batch + j: by thread pt, or for the
        // learned traveltime (the larger batches) by wave 0 alone, lane pt % 64
        const int n_pairs = kNPil * P.batch;
        const unsigned m_batch = div_magic(P.batch);
        const int pt_p0 = tid / P.batch, pt_j0 = tid % P.batch;

        // Learned traveltime on the t0 lattice.  The network's outputs g (the
        // legs' learned factors) are evaluated at the lattice nodes for three
        // pillars of the workgroup's group of kNPil (0, p_mid and the last
        // valid one, p_hi), quadratic in the pillar index in between (a leg's
        // squared distance is quadratic in it: a line through two pillars
        // would bias every time the same way); along t0 each slot
        // takes g from the cubic Hermite polynomial of its segment [n_j,
        // n_j+1] of the axis lattice (the segment clamped to the window's first
        // and last), through g at the two nodes with slopes from differences
        // over nodes j - 1 .. j + 2; the leg times are then formed per sample
        // with the sample's own velocity and eta.  The window's nodes are jf ..
        // jl; with one neighbour on each side the item tables "extended" nodes
        // jf - 1 .. jl + 1 (local index i = j - jf + 1, nx = nn + 2 of them;
        // nodes off the axis are never evaluated).  Per (pillar, extended
        // node): the smallest taper slope (widest aperture) over the samples
        // whose segment reads the node (segments j - 2 .. j + 1): a node no
        // pillar of the group sees inside its aperture needs no network
        // evaluation; and the node's network inputs (t0 / 2, slowness^2 or
        // vg^2, encoded velocity and eta / gradient), so that the evaluation
        // reads LDS only.
        int nidx[MW];
        float nfr[MW];
        int nn = 0;  // lattice nodes of this window (wave-uniform)
        int jf = 0;
        if (learned_tt(TT)) {
            const int* npos = reinterpret_cast<const int*>(smem + P.lat_pos_off);
            auto upper = [&](int k) {  // first node past sample k
                int lo = 0, hi = P.lat_total;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (npos[mid] <= k) lo = mid + 1;
                    else hi = mid;
                }
                return lo;
            };
            const int kend = min(kbase + win, P.nt);
            jf = __builtin_amdgcn_readfirstlane(min(upper(kbase) - 1, P.lat_total - 2));
            const int jl = max(upper(kend - 2), jf + 1);
            nn = __builtin_amdgcn_readfirstlane(jl - jf + 1);
#pragma unroll
            for (int m = 0; m < MW; ++m) {
                const int k = min(kbase + 64 * (tw + kTW * m) + lane, kend - 1);
                const int j = max(jf, min(upper(k) - 1, jl - 1));
                nidx[m] = j - jf;  // segment of the window
                nfr[m] = (float)(k - npos[j]) / (float)(npos[j + 1] - npos[j]);
            }
            const int nx = nn + 2;
            float* cimin = reinterpret_cast<float*>(smem + P.lat_cimin_off);
            float4* ntab = reinterpret_cast<float4*>(smem + P.node_off);
            for (int e = tid; e < kNPil * nx; e += kThreads) {
                const int p = e / nx, i = e - p * nx;
                const int pil = pil0 + p;
                const int j = jf - 1 + i;
                // window segments that read node j: j - 2 .. j + 1
                const int sa = max(j - 2, jf), sb = min(j + 1, jl - 1);
                float c = INFINITY;
                float4 nd = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (pil < npil && j >= 0 && j < P.lat_total && sa <= sb) {
                    const int k0 = sa == jf ? kbase : npos[sa];
                    const int k1 = sb == jl - 1 ? kend : npos[sb + 1];
                    for (int k = k0; k < k1; ++k) c = fminf(c, T.kc[(size_t)pil * T.kc_pstride + k].z);
                    const int kk = npos[j];
                    const float tau = 0.5f * P.t0k[kk];
                    float s, v, eta;
                    if (TT == PSTM_TT_VTI_LEARNED) {
                        s = T.kc[(size_t)pil * T.kc_pstride + kk].x;  // 1/v^2
                        v = frsq(s);
                        eta = 0.5f * T.ek[(size_t)pil * T.ek_pstride + kk].x;
                    } else {
                        s = P.qk[kk];  // vg^2
                        v = fsqrt(s);
                        eta = (float)P.ck;
                    }
                    nd = make_float4(tau, s, fmaf(v - P.m_v_lo, P.m_v_sc, -1.0f), fmaf(eta, P.m_eta_sc, -1.0f));
                }
                cimin[e] = c;
                ntab[e] = nd;
            }
        }

        // Pipelined staging (pipe, on request; never for the learned
        // traveltime, whose rows go by direct loads so that no registers are
        // held across the network evaluation): while a batch computes, each thread holds
        // up to kPF dwords of the next batch's rows in flight in registers; they
        // are written to LDS after the batch's last barrier.
        uint32_t pf[kPF];
        const int rl_dw = P.rowlen * (int)(sizeof(Row) / 4);
        // row jrow of the tile's list (r-th of its batch), its first dword in
        // global memory: the block row from s_rownext[rpar] (written by the
        // batch before; for the first batch at the item start), or the tile's
        // own order for AA rows read in place.  The LDS array is indexed by
        // name so its reads stay LDS reads: through a plain pointer they would
        // be flat loads, whose wait also waits for the row loads in flight
        auto row_dw = [&](int jrow, int rpar, int r) -> const uint32_t* {
            if (AA && !P.prow_global) return reinterpret_cast<const uint32_t*>(P.prow + (size_t)jrow * P.rowlen);
            const int rw = s_rownext[rpar][r];
            if constexpr (AA) return reinterpret_cast<const uint32_t*>(P.prow + (size_t)rw * P.rowlen);
            else return reinterpret_cast<const uint32_t*>(P.amp + (size_t)rw * P.ns);
        };
        auto prefetch = [&](int jb, int rpar) {
            const int nbb = min(P.batch, T.ntr - jb);
            if (nbb <= 0) return;
            const int n = nbb * rl_dw;
            int r = 0, c = tid;
            while (c >= rl_dw) {
                c -= rl_dw;
                ++r;
            }
            const uint32_t* src = row_dw(jb + min(r, nbb - 1), rpar, min(r, nbb - 1));
#pragma unroll
            for (int i = 0; i < kPF; ++i) {
                if (tid + i * kThreads < n) pf[i] = src[c];
                c += kThreads;
                bool moved = false;
                while (c >= rl_dw) {
                    c -= rl_dw;
                    ++r;
                    moved = true;
                }
                if (moved && r < nbb) src = row_dw(jb + r, rpar, r);
            }
        };
        // Rows by direct global -> LDS loads (every unpipelined batch; pipelined
        // AA rows too), into the second of two row buffers when pipelined, else
        // into the one buffer: (row, 64-dword chunk) units dealt round-robin
        // over waves w0.., one wave instruction each (wave-uniform LDS
        // destination, one dword per lane).  The batch's block rows are read
        // from LDS once per call and the unit's row taken from lane r, so
        // that the loop holds no memory read: a read whose result the loop
        // waited for would also wait for every row load issued before it.
        int cur = 0;
        auto dma = [&](int jb, int buf, int rpar, int w0) {
            {
                const int nbb = min(P.batch, T.ntr - jb);
                if (nbb <= 0 || wave < w0) return;
                const int nw = kWaves - w0;
                const int nck = (rl_dw + 63) >> 6;
                uint32_t* dst0 = reinterpret_cast<uint32_t*>(smem + (size_t)buf * P.rows_bytes);
                const bool listed = !(AA && !P.prow_global);
                const int lrow = listed && lane < nbb ? s_rownext[rpar][lane] : jb + lane;
                int r = 0, ck = wave - w0;
                while (ck >= nck) {
                    ck -= nck;
                    ++r;
                }
                while (r < nbb) {
                    const int rw = __builtin_amdgcn_readlane(lrow, r);
                    const uint32_t* src;
                    if constexpr (AA) src = reinterpret_cast<const uint32_t*>(P.prow + (size_t)rw * P.rowlen);
                    else src = reinterpret_cast<const uint32_t*>(P.amp + (size_t)rw * P.ns);
                    const int col = ck * 64 + lane;
                    if (col < rl_dw) __builtin_amdgcn_global_load_lds(src + col, dst0 + r * rl_dw + ck * 64, 4, 0, 0);
                    ck += nw;
                    while (ck >= nck) {
                        ck -= nck;
                        ++r;
                    }
                }
            }
        };
        // the first batch's block rows where every later batch finds them
        // (s_rownext[1], parity of the batch before the first)
        if (tid < P.batch) s_rownext[1][tid] = tid < T.ntr ? (T.tidx ? T.tidx[tid] : tid) : 0;
        if (tid == 0) s_any[0] = s_any[1] = -1;
        if (learned_tt(TT) && tid < kNPil) {
            float w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
            const float fp = (float)tid;
            if (p_hi >= 2) {
                const float b = (float)p_mid, c = (float)p_hi;
                w0 = (fp - b) * (fp - c) / (b * c);
                w1 = fp * (fp - c) / (b * (b - c));
                w2 = fp * (fp - b) / (c * (c - b));
            } else {
                w2 = p_hi == 1 ? fp : 0.0f;
                w0 = 1.0f - w2;
            }
            s_lw[tid] = make_float4(w0, w1, w2, 0.0f);
        }
        __syncthreads();
        if (pipe) {
            if (AA) dma(0, 0, 1, 0);
            else prefetch(0, 1);
        }

        // Unpipelined staging (early_dma, the default): the batch's rows (f32,
        // or the f64 AA rows) go by direct global -> LDS loads into the one
        // row buffer, issued at the batch start
        // so that they are in flight while the pair table is built: every load
        // of a wave in flight at once and no registers held (a loop of
        // load-then-store through registers waits out the memory latency once
        // per element).  The row indices come from s_rownext, written by the
        // batch before.  The waves that build the pair table issue none: their
        // own global reads would wait behind the row loads.
        const bool early_dma = !pipe;
        const int w_pairs = learned_tt(TT) ? 1 : (n_pairs + 63) / 64;  // waves holding the pair table
        // Pair inputs one batch ahead (w_pairs == 1): wave 0 loads batch jb's
        // geometry and midpoints (rows from s_rownext[rpar]) and batch jb +
        // batch's block rows into s_stage / s_stage_row during the batch
        // before (after that batch's wait for its row loads, or, skipped,
        // after its pair-table barrier), so that the pair table (on which
        // every wave waits) reads LDS only; the loads land during the trace
        // loop.  Only wave 0 reads them, after its own vmcnt wait.
        const bool staged_pairs = w_pairs == 1;
        auto stage = [&](int jb, int rpar) {
            const int nbb = min(P.batch, T.ntr - jb);
            if (nbb <= 0) return;
            for (int c = 0; c * 64 < nbb * 12; ++c) {
                const int d = c * 64 + lane;
                if (d < nbb * 12) {
                    const int jt = d / 12, f = d - jt * 12;
                    const uint32_t* src;
                    if (f < 8) {
                        src = reinterpret_cast<const uint32_t*>(T.geo + 2 * (jb + jt)) + f;
                    } else {
                        const int trow = s_rownext[rpar][jt];
                        src = reinterpret_cast<const uint32_t*>((f < 10 ? P.mx64 : P.my64) + trow) + (f & 1);
                    }
                    __builtin_amdgcn_global_load_lds(src, s_stage + c * 64, 4, 0, 0);
                }
            }
            const int jn = jb + P.batch + lane;
            if (T.tidx && lane < P.batch && jn < T.ntr) __builtin_amdgcn_global_load_lds(T.tidx + jn, s_stage_row, 4, 0, 0);
        };
        for (int j0 = 0, par = 0; j0 < T.ntr; j0 += P.batch, par ^= 1) {
            const int nb = min(P.batch, T.ntr - j0);
            const bool staged = staged_pairs && j0 > 0;
            if (staged && wave == 0) {
                // this wave's staging loads (issued a batch ago) have landed;
                // no LDS read of them may move above the wait
                __builtin_amdgcn_s_waitcnt(0x0F70);
                asm volatile("" ::: "memory");
            }
            if (early_dma) dma(j0, 0, par ^ 1, w_pairs);
            // pipelined f32 rows, loaded into registers during the batch
            // before (or before the first): to LDS at the batch start, while
            // the row buffer is free (the batch before ended with a barrier),
            // so that those registers are free for the rest of the batch (the
            // network evaluation above all); visible after the pair-table
            // barrier.  A skipped batch stores rows nobody reads.
            if (!AA && pipe) {
                uint32_t* dst = reinterpret_cast<uint32_t*>(rows);
                const int n = nb * rl_dw;
#pragma unroll
                for (int i = 0; i < kPF; ++i)
                    if (tid + i * kThreads < n) dst[tid + i * kThreads] = pf[i];
            }
            // (pillar, trace) pair table; the max-aperture test is the
            // oracle's f64 one (numba_cpu_optimized.py:167-173)
            int inap = 0;
            if (learned_tt(TT) && tid == 0) s_nrows = 0;
            if (tid < P.batch) {  // wave 0, which builds pairs and issues no row loads
                // the next batch's rows: its row loads and its pair table read
                // them from here
                const int jn = j0 + P.batch + tid;
                s_rownext[par][tid] = jn < T.ntr ? (T.tidx ? (staged ? s_stage_row[tid] : T.tidx[jn]) : jn) : 0;
            }
            auto build_pair = [&](int pt_p, int pt_j) {
                const int pil = pil0 + pt_p;
                float4 e = make_float4(0.0f, 0.0f, INFINITY, 0.0f);
                float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (pt_j < nb && pil < npil) {
                    // staged inputs from LDS; else every global load of the pair
                    // issued at once (one memory latency), the block row from
                    // the previous batch's table
                    float4 g0, g1;
                    double mx64, my64;
                    if (staged) {
                        const float4* s4 = reinterpret_cast<const float4*>(s_stage) + 3 * pt_j;
                        const double* s8 = reinterpret_cast<const double*>(s_stage) + 6 * pt_j;
                        g0 = s4[0];
                        g1 = s4[1];
                        mx64 = s8[4];
                        my64 = s8[5];
                    } else {
                        const int trow = s_rownext[par ^ 1][pt_j];
                        g0 = T.geo[2 * (j0 + pt_j)];
                        g1 = T.geo[2 * (j0 + pt_j) + 1];
                        mx64 = P.mx64[trow];
                        my64 = P.my64[trow];
                    }
                    const double pox64 = T.xax[pil / T.ny], poy64 = T.yax[pil % T.ny];
                    const float pox = (float)(pox64 - T.xax[0]), poy = (float)(poy64 - T.yax[0]);
                    const double dx64 = __dsub_rn(pox64, mx64);
                    const double dy64 = __dsub_rn(poy64, my64);
                    const double dm64 = dist64(dx64, dy64);
                    // learned traveltimes: the leg terms of every pillar, also
                    // outside the aperture (an end pillar's network inputs)
                    const bool in_ap64 = !(dm64 > P.max_ap64);
                    if (in_ap64 || learned_tt(TT)) {
                        if (in_ap64) inap = 1;
                        const float dxs = pox - g0.x, dys = poy - g0.y;
                        const float dxr = pox - g0.z, dyr = poy - g0.w;
                        const float dxm = pox - g1.x, dym = poy - g1.y;
                        const float dm = sqrtf(dxm * dxm + dym * dym);
                        const float ds2 = dxs * dxs + dys * dys;
                        const float dr2 = dxr * dxr + dyr * dyr;
                        if (TT == PSTM_TT_CURVED) {
                            const float a_s = P.ck2 * ds2, a_r = P.ck2 * dr2;
                            const float den_s = P.v0 + sqrtf(P.v0 * P.v0 + a_s);
                            const float den_r = P.v0 + sqrtf(P.v0 * P.v0 + a_r);
                            e = make_float4(a_s, a_r, dm, g1.z);
                            f = make_float4(1.0f / (den_s * den_r), 0.0f, 0.0f, 0.0f);
                        } else if (TT == PSTM_TT_CURVED_LEARNED) {
                            e = make_float4(P.ck2 * ds2, P.ck2 * dr2, in_ap64 ? dm : INFINITY, g1.z);
                        } else if (TT == PSTM_TT_VTI_LEARNED) {
                            e = make_float4(ds2, dr2, in_ap64 ? dm : INFINITY, g1.z);
                        } else {
                            e = make_float4(ds2, dr2, dm, g1.z);
                            // AA: offsets pre-scaled by bin size / dt, so the operator
                            // dip in samples along x is cs * (f.x y1 + f.z y2)
                            f = make_float4(dxs * P.gdx_idt, dys * P.gdy_idt, dxr * P.gdx_idt, dyr * P.gdy_idt);
                        }
                    }
                }
                pairs[(pt_p * P.batch + pt_j) * kPairVec] = e;
                if (kPairVec == 2) pairs[(pt_p * P.batch + pt_j) * kPairVec + 1] = f;
            };
            if constexpr (learned_tt(TT)) {
                if (wave == 0) {
                    for (int pt = lane; pt < n_pairs; pt += 64) {
                        const int pp = div_by(pt, m_batch, P.batch);
                        build_pair(pp, pt - pp * P.batch);
                    }
                }
            } else if (tid < n_pairs) {
                build_pair(pt_p0, pt_j0);
            }
            // the written j0 tells this batch from the last one of its parity
            if (wave_any(inap) && lane == 0) s_any[par] = j0;
            lds_barrier();
            const bool any_inap = __builtin_amdgcn_readfirstlane(s_any[par]) == j0;
            // the next batch's pair inputs: issued after this batch's last wait
            // for row loads (a wait would also wait for them)
            if (!any_inap) {
                if (staged_pairs && wave == 0) stage(j0 + P.batch, par);
                if (pipe) {  // this batch's rows are not needed
                    if (AA) {
                        dma(j0 + P.batch, cur ^ 1, par, 0);
                        cur ^= 1;
                    } else {
                        prefetch(j0 + P.batch, par);
                    }
                }
                continue;
            }

            if (learned_tt(TT)) {
                // network on the t0 lattice at the group's three evaluated
                // pillars (pe = 0, 1, 2: pillars 0, p_mid, p_hi; pe = 1 only for
                // groups of three or more, pe = 2 for two or more): the rows
                // (pe, trace, extended node) whose node some pillar of the
                // group reads inside its time-dependent aperture are listed (in
                // any order), the others marked missing (NaN); then 16 listed
                // rows per wavefront step
                const int nx = nn + 2;
                const int R = 3 * P.batch * nx;
                const unsigned m_bnx = div_magic(P.batch * nx), m_nx = div_magic(nx);
                float2* lat = reinterpret_cast<float2*>(smem + P.lat_off);
                const float* cimin = reinterpret_cast<const float*>(smem + P.lat_cimin_off);
                unsigned short* rlist = reinterpret_cast<unsigned short*>(smem + P.lat_list_off);
                for (int r0 = 0; r0 < R; r0 += kThreads) {
                    const int r = r0 + tid;
                    bool valid = false;
                    if (r < R) {
                        const int pe = div_by(r, m_bnx, P.batch * nx), jn = r - pe * (P.batch * nx);
                        const int jt = div_by(jn, m_nx, nx), node = jn - jt * nx;
                        if (pe == 0 || (pe == 1 && p_hi >= 2) || (pe == 2 && p_hi >= 1)) {
                            for (int p = 0; p <= p_hi && !valid; ++p) {
                                const float ez = pairs[(p * P.batch + jt) * kPairVec].z;
                                valid = ez != INFINITY && fmaf(ez, cimin[p * nx + node], P.taper_off) <= 1.0f + P.eps_x;
                            }
                        }
                        if (!valid) lat[r] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
                    }
                    const unsigned long long bal = __ballot(valid);
                    int base = 0;
                    if (lane == 0 && bal) base = atomicAdd(&s_nrows, (int)__popcll(bal));
                    base = __shfl(base, 0, 64);
                    if (valid) rlist[base + (int)__popcll(bal & ((1ull << lane) - 1ull))] = (unsigned short)r;
                }
                lds_barrier();
                const int nrows = s_nrows;
                const float4* ntab = reinterpret_cast<const float4*>(smem + P.node_off);
                LdsA* wa = (LdsA*)(smem + P.mlp_off);
                LdsB* wb = (LdsB*)(smem + P.mlp_off + mlp_nv * 16);
                for (int c0 = wave * 16; c0 < nrows; c0 += kWaves * 16) {
                    const int li = c0 + (lane & 15);
                    const bool valid = li < nrows;
                    const int r = valid ? (int)rlist[li] : 0;
                    float x[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                    if (valid) {
                        const int pe = div_by(r, m_bnx, P.batch * nx), jn = r - pe * (P.batch * nx);
                        const int jt = div_by(jn, m_nx, nx), node = jn - jt * nx;
                        const int pend = pe == 0 ? 0 : (pe == 1 ? p_mid : p_hi);
                        const float4 e = pairs[(pend * P.batch + jt) * kPairVec];
                        const float4 nd = ntab[pend * nx + node];  // {tau, s, v enc., eta enc.}
                        const float tau = nd.x;
                        float sn_s, sn_r;
                        if (TT == PSTM_TT_VTI_LEARNED) {
                            const float q = tau * tau;
                            const float hs = e.x * nd.y, hr = e.y * nd.y;
                            const float ts2 = q + hs, tr2 = q + hr;
                            sn_s = ts2 > 0.0f ? fsqrt(hs * frcp(ts2)) : 0.0f;
                            sn_r = tr2 > 0.0f ? fsqrt(hr * frcp(tr2)) : 0.0f;
                        } else {
                            // curved: sines k h / sqrt(vg^2 + k^2 h^2)
                            sn_s = fsqrt(e.x * frcp(e.x + nd.y));
                            sn_r = fsqrt(e.y * frcp(e.y + nd.y));
                        }
                        x[0] = fmaf(tau, P.m_tau_sc, -1.0f);
                        x[1] = fmaf(2.0f, sn_s, -1.0f);
                        x[2] = fmaf(2.0f, sn_r, -1.0f);
                        x[3] = nd.z;
                        x[4] = nd.w;
                    }
                    float y0, y1;
                    if (P.mlp_split)
                        mlp_eval16<true>(wa, wb, x, y0, y1);
                    else
                        mlp_eval16<false>(wa, wb, x, y0, y1);
                    if (lane < 16 && valid) lat[r] = make_float2(y0 * P.m_inv_out, y1 * P.m_inv_out);
                }
            }

            // stage the batch of trace rows in LDS (coalesced)
            if (AA && pipe) {
                rows = reinterpret_cast<Row*>(smem + (size_t)cur * P.rows_bytes);
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's row loads have landed
            } else if (pipe) {
                // f32 rows: stored at the batch start
            } else {
                // the batch's rows, loaded at the batch start (early_dma)
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's row loads have landed
            }
            // pipelined f32 rows are visible since the pair-table barrier: only
            // DMA'd or loaded rows and the lattice need this one
            if (AA || !pipe || learned_tt(TT)) lds_barrier();
            if (staged_pairs && wave == 0) stage(j0 + P.batch, par);
            if (pipe) {  // in flight during the batch
                if (AA) dma(j0 + P.batch, cur ^ 1, par, 0);
                else prefetch(j0 + P.batch, par);
            }
            if (learned_tt(TT)) {
                // per (pillar, trace, window segment) the cubic Hermite
                // coefficients of both legs' g in the segment's fraction f,
                // g = a + f (b + f (c + f d)): g at the four nodes j - 1 .. j + 2
                // from the three evaluated pillars (Lagrange weights in the
                // pillar index: quadratic; a line for a group of two, the value
                // itself for one), values at n_j and n_j+1 with slopes by
                // node_derivative; a segment with a node not evaluated is read
                // by no slot inside the aperture (zeros)
                const int nx = nn + 2, nseg = nn - 1;
                const int* npos = reinterpret_cast<const int*>(smem + P.lat_pos_off);
                const float2* lat = reinterpret_cast<const float2*>(smem + P.lat_off);
                float4* coef = reinterpret_cast<float4*>(smem + P.coef_off);
                const unsigned m_nseg = div_magic(nseg), m_b = div_magic(P.batch);
                for (int e = tid; e < kNPil * P.batch * nseg; e += kThreads) {
                    const int pp = div_by(e, m_nseg, nseg), sg = e - pp * nseg;
                    const int p = div_by(pp, m_b, P.batch), jt = pp - p * P.batch;
                    const int j = jf + sg;
                    float4 cs4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), cr4 = cs4;
                    if (p <= p_hi) {
                        // weights of pillars 0, p_mid, p_hi for pillar p
                        const float4 lw = s_lw[p];
                        const float w0 = lw.x, w1 = lw.y, w2 = lw.z;
                        const float2* ra = lat + jt * nx + sg;                  // pe 0, nodes j - 1 .. j + 2
                        const float2* rb = lat + (P.batch + jt) * nx + sg;      // pe 1
                        const float2* rc = lat + (2 * P.batch + jt) * nx + sg;  // pe 2
                        float gs[4], gr[4], x[4];
                        bool ok[4];
#pragma unroll
                        for (int n = 0; n < 4; ++n) {
                            const int jn = j - 1 + n;
                            const float2 ga = ra[n];
                            const float2 gb = p_hi >= 2 ? rb[n] : make_float2(0.0f, 0.0f);
                            const float2 gc = p_hi >= 1 ? rc[n] : make_float2(0.0f, 0.0f);
                            ok[n] = !__builtin_isnan(ga.x) && !__builtin_isnan(gb.x) && !__builtin_isnan(gc.x);
                            gs[n] = fmaf(w0, ga.x, fmaf(w1, gb.x, w2 * gc.x));
                            gr[n] = fmaf(w0, ga.y, fmaf(w1, gb.y, w2 * gc.y));
                            x[n] = (float)npos[min(max(jn, 0), P.lat_total - 1)] +
                                   (float)(jn < 0 ? jn : (jn >= P.lat_total ? jn - P.lat_total + 1 : 0));
                        }
                        if (ok[1] && ok[2]) {
                            const float L = x[2] - x[1];
                            auto herm = [&](const float (&g)[4]) {
                                const float b0 = L * node_derivative(g, ok, x, 1), b1 = L * node_derivative(g, ok, x, 2);
                                return make_float4(g[1], b0, 3.0f * (g[2] - g[1]) - 2.0f * b0 - b1,
                                                   2.0f * (g[1] - g[2]) + b0 + b1);
                            };
                            cs4 = herm(gs);
                            cr4 = herm(gr);
                        }
                    }
                    coef[2 * e] = cs4;
                    coef[2 * e + 1] = cr4;
                }
                lds_barrier();
            }

            // the traces of the batch; EARLY: this wave holds t0 <= 2 ms samples,
            // where t <= 1 ms can switch the spreading/obliquity weights off
            // (wave-uniform, rare: one copy of the loop without that test)
            auto trace_loop = [&](auto early_c) {
                constexpr bool EARLY = decltype(early_c)::value;
                // the sample offset held in a vector register: the sample position
                // fmaf(t, idt, uoff) then needs no copy of a scalar operand per slot
                float uoff_v = P.uoff;
                asm("" : "+v"(uoff_v));
                for (int jj = 0; jj < nb; ++jj) {
                    const Row* row = rows + (size_t)jj * P.rowlen;
                    bool any_near = false;  // some slot within eps of a threshold (rare; wave-uniform)
                    // interpolators that jump at fraction 0.5 have a near lane in
                    // most traces: the slots (bit p * MW + m) holding one, so
                    // that only those are recomputed
                    unsigned near_slots = 0;
                    // both pairs' records read at once: one LDS round trip per trace
                    float4 ep[kPW];
    #pragma unroll
                    for (int p = 0; p < kPW; ++p) ep[p] = pairs[((pg * kPW + p) * P.batch + jj) * kPairVec];
    #pragma unroll
                    for (int p = 0; p < kPW; ++p) {
                        const int pi = (pg * kPW + p) * P.batch + jj;
                        const float4 e = ep[p];
                        if (e.z == INFINITY) continue;  // pair outside max aperture (wave-uniform)
                        float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (kPairVec == 2) f = pairs[pi * kPairVec + 1];
                        // Every slot of the pair is evaluated without per-slot branches so
                        // that the MW independent chains (sqrt, LDS reads) interleave;
                        // slots outside the aperture or the trace read a safe sample and
                        // add zero.
                        // blk[m]: some lane of sample block m is inside the
                        // time-dependent aperture (wave-uniform); the aperture grows
                        // with t0, so the early blocks of distant pairs are skipped
                        float xt[MW];
                        bool blk[MW];
                        bool any_in = false;
    #pragma unroll
                        for (int m = 0; m < MW; ++m) {
                            xt[m] = fmaf(e.z, ci[p][m], P.taper_off);
                            // !(x > th), not x <= th: a predicate of its own, so that the
                            // compiler does not carry this compare's lane mask into the
                            // slot's block (through a VGPR) in place of in_ap's compare
                            blk[m] = wave_any(!(xt[m] > ((COUNT && !learned_tt(TT)) ? 1.0f + P.eps_x : 1.0f)));
                            any_in |= blk[m];
                        }
                        if (!any_in) continue;  // time-dependent aperture excludes every slot
                        // learned: the pair's Hermite coefficients of g per window segment
                        const float4* cfp = reinterpret_cast<const float4*>(smem + P.coef_off) + (size_t)pi * 2 * (nn - 1);
                        // one pass per sample block, so that only one block's
                        // temporaries are live at a time (fewer registers)
    #pragma unroll
                        for (int m = 0; m < MW; ++m) {
                            if (!blk[m]) continue;
                            float y1, y2;
                            if constexpr (learned_tt(TT)) {
                                // both legs' g: the segment's cubics in the slot's fraction
                                const float4 ca = cfp[2 * nidx[m]], cb = cfp[2 * nidx[m] + 1];
                                const float fr = nfr[m];
                                e2[p][m] = fmaf(fmaf(fmaf(ca.w, fr, ca.z), fr, ca.y), fr, ca.x);
                                e1[p][m] = fmaf(fmaf(fmaf(cb.w, fr, cb.z), fr, cb.y), fr, cb.x);
                            }
                            // (t0 <= 0 only in a block whose first t0 is <= 2 ms: early)
                            const float t = contrib_time<TT, AA>(P, e, f, q[m], cs[p][m], e2[p][m], e1[p][m], t0r[m],
                                                                 etar[p][m], y1, y2, EARLY && ((early >> m) & 1u));
                            const float u = fmaf(t, P.idt, uoff_v);
                            // u < umax - eps: inside; [umax - eps, umax + eps): decided in f64
                            const bool in_ap = xt[m] <= ((COUNT && !learned_tt(TT)) ? 1.0f - P.eps_x : 1.0f);
                            const bool below = u < P.umax_lo;
                            // thresholds that can hold lanes below umax - eps (the learned
                            // time has no f64 counterpart: none)
                            bool near_o = false;
                            if (LOB && !learned_tt(TT)) near_o = near_o || (in_ap && u < P.lo_eps && u > -P.lo_eps);
                            if (COUNT && !learned_tt(TT)) near_o = near_o || (fabsf(xt[m] - 1.0f) < P.eps_x);
                            // nearest sample and windowed sinc jump at fraction 0.5 (the
                            // sample / the tap window moves by one): decided in f64 there
                            if (kHalfJump<INTERP>)
                                near_o = near_o || (in_ap && below &&
                                                     fabsf(__builtin_amdgcn_fractf(u) - 0.5f) < fmaf(u, P.eps_half, 1e-5f));
                            // some lane of the slot near a threshold (wave-uniform): the
                            // trace-end band [umax - eps, umax + eps) inside the aperture,
                            // a scalar test on the three compare masks, or near_o; the
                            // learned time has no f64 counterpart to decide against
                            bool near_w = false;
                            if (!learned_tt(TT)) {
                                // (!(u < lo) as a compare of its own: the complement of below's
                                // mask would be built through a VGPR)
                                unsigned long long band =
                                    wave_mask(in_ap) & wave_mask(!(u < P.umax_lo)) & wave_mask(u < P.umax_hi);
                                // taken here, while the masks are in scalar registers (sunk
                                // past the taper branch, below's mask would be carried
                                // across it through a VGPR: two more VALU operations)
                                asm volatile("" : "+s"(band));
                                near_w = band != 0;
                                if (LOB || COUNT || kHalfJump<INTERP>) near_w = near_w || wave_any(near_o);
                            }
                            any_near = any_near || near_w;
                            if constexpr (kHalfJump<INTERP>) near_slots |= near_w ? 1u << (p * MW + m) : 0u;
                            // band lanes are never below umax - eps: only near_o takes lanes
                            // out of the ones below
                            bool ok = in_ap && (learned_tt(TT) ? u < P.umax : below);
                            if ((LOB || COUNT || kHalfJump<INTERP>) && !learned_tt(TT)) ok = ok && !near_o;
                            if (LOB) ok = ok && u >= 0.0f;
                            const float us = ok ? u : 0.0f;
                            const int i0 = (int)us;
                            const float f0 = __builtin_amdgcn_fractf(us);
                            // the row reads are issued before the weight is worked out
                            AaTaps ta;
                            LinTaps tl;
                            if constexpr (AA) ta = aa_fetch(P, row, i0, f, cs[p][m], y1, y2, s_invsq);
                            else if constexpr (INTERP == PSTM_INTERP_LINEAR) tl = lin_fetch(row, i0);
                            float w = ok ? e.w : 0.0f;
                            // taper zone 0 < x_t <= 1 on some lane (wave-uniform test: the
                            // ballots of two plain compares, ANDed as masks, stay scalar)
                            if ((wave_mask(xt[m] > 0.0f) & wave_mask(xt[m] <= 1.0f)) != 0) {
                                // kept a branch (most waves skip it): the empty asm stops
                                // the compiler from computing the taper for every wave
                                asm volatile("");
                                // a select inside: the taper runs for the whole wave,
                                // without exec-mask juggling
                                const float tw = taper_weight(__builtin_amdgcn_fmed3f(xt[m], 0.0f, 1.0f));
                                w = xt[m] > 0.0f ? w * tw : w;
                            }
                            float amp;
                            if constexpr (AA) amp = aa_finish(ta, f0);
                            else if constexpr (INTERP == PSTM_INTERP_LINEAR) amp = lin_finish(tl, f0);
                            else amp = interp_at<INTERP>(row, i0, f0, P.ns);
                            if (W2 || P.wpow != 0) {
                                // spreading / obliquity: cw / t^wpow; cw, constant over the
                                // traces, is applied to the slot's total when the item ends,
                                // except in early blocks, where t <= 1 ms switches the factor off
                                const float r = frcp(t);
                                float wf = (W2 || P.wpow == 2) ? r * r : r;
                                // t > 1 ms is certain unless the block holds t0 <= 2 ms (early)
                                if (EARLY && (early & (1u << m))) wf = t > 0.001f ? cw_at(p, m) * wf : 1.0f;
                                w *= wf;
                            }
                            acc[p][m] = fmaf(amp, w, acc[p][m]);
                            if (COUNT) cnt += ok ? 1 : 0;
                        }
                    }
                    // slots near a threshold: recompute which, decide in f64 as the
                    // oracle does (rare; the slots were left out above)
                    if (!learned_tt(TT) && any_near) {
    #pragma unroll
                        for (int p = 0; p < kPW; ++p) {
                            const int pi = (pg * kPW + p) * P.batch + jj;
                            const float4 e = pairs[pi * kPairVec];
                            if (e.z == INFINITY) continue;
                            float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
                            if (kPairVec == 2) f = pairs[pi * kPairVec + 1];
    #pragma unroll
                            for (int m = 0; m < MW; ++m) {
                                if (kHalfJump<INTERP> && !(near_slots & (1u << (p * MW + m)))) continue;
                                const float xt = fmaf(e.z, ci[p][m], P.taper_off);
                                float y1, y2;
                                const float t = contrib_time<TT, AA>(P, e, f, q[m], cs[p][m], e2[p][m], e1[p][m],
                                                                     t0r[m], etar[p][m], y1, y2);
                                const float u = fmaf(t, P.idt, P.uoff);
                                const bool in_ap = xt <= (COUNT ? 1.0f - P.eps_x : 1.0f);
                                bool near = in_ap && u >= P.umax_lo && u < P.umax_hi;
                                if (LOB) near = near || (in_ap && u < P.lo_eps && u > -P.lo_eps);
                                if (COUNT) near = near || (fabsf(xt - 1.0f) < P.eps_x);
                                if (kHalfJump<INTERP>)
                                    near = near || (in_ap && u < P.umax_lo &&
                                                    fabsf(__builtin_amdgcn_fractf(u) - 0.5f) < fmaf(u, P.eps_half, 1e-5f));
                                if (!near) continue;
                                const int pil = pil0 + pg * kPW + p;
                                const int k = kbase + 64 * (tw + kTW * m) + lane;
                                const int trow = T.tidx ? T.tidx[j0 + jj] : j0 + jj;
                                double u64;
                                if (exact_decision<TT>(P, T, pil, trow, k, &u64)) {
                                    // the nearest sample is rounded in f64: (int)(u + 0.5)
                                    const int i0 = INTERP == PSTM_INTERP_NEAREST ? (int)(u64 + 0.5) : (int)u64;
                                    const float f0 = INTERP == PSTM_INTERP_NEAREST ? 0.0f : (float)(u64 - (double)i0);
                                    // cw in early blocks only (applied to the totals elsewhere)
                                    const float cwe = (early & (1u << m)) ? cw_at(p, m) : 1.0f;
                                    acc[p][m] = fmaf(contrib_amp<AA, INTERP>(P, row, i0, f0, f, cs[p][m], y1, y2, s_invsq),
                                                     contrib_weight(P, e.w, xt, t, cwe), acc[p][m]);
                                    if (COUNT) ++cnt;
                                }
                            }
                        }
                    }
                }
            };
            if (early) trace_loop(std::true_type{});
            else trace_loop(std::false_type{});
#pragma unroll
            for (int p = 0; p < kPW; ++p)
#pragma unroll
                for (int m = 0; m < MW; ++m) {
                    if (learned_tt(TT)) totf[p][m] += acc[p][m];
                    else tot[(p * MW + m) * kThreads + tid] += (double)acc[p][m];
                    acc[p][m] = 0.0f;
                }
            lds_barrier();
            if (AA && pipe) cur ^= 1;
        }
        // a skipped last batch leaves its row loads in flight: none may outlive
        // the item (nor the workgroup, whose LDS is then reassigned)
        if (early_dma || staged_pairs) __builtin_amdgcn_s_waitcnt(0x0F70);

        // the workgroup is the only writer of these cells
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            const int k = kbase + 64 * (tw + kTW * m) + lane;
            if (k >= P.nt) continue;
#pragma unroll
            for (int p = 0; p < kPW; ++p) {
                const int pil = pil0 + pg * kPW + p;
                if (pil < npil) {
                    // the total times cw (early blocks carry it per contribution)
                    const double cwb = (early & (1u << m)) ? 1.0 : (double)cw_at(p, m);
                    float* dst = T.image + (size_t)pil * P.nt + k;
                    const double sum = learned_tt(TT) ? (double)totf[p][m]
                                                      : tot[(p * MW + m) * kThreads + tid];
                    *dst = *dst + (float)(sum * cwb);
                }
            }
        }
        __syncthreads();
    }

    if (COUNT) {
        __shared__ unsigned long long red[kWaves];
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
        if (lane == 0) red[wave] = cnt;
        __syncthreads();
        if (tid == 0) {
            unsigned long long s = 0;
            for (int w = 0; w < kWaves; ++w) s += red[w];
            P.cnt_part[blockIdx.x] = s;
        }
    }
}


#ifndef PSTM_GATHER_PART
__global__ void k_sum_counts(const unsigned long long* part, int n, unsigned long long* out) {
    __shared__ unsigned long long red[256];
    unsigned long long s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out += red[0];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Fixed scratch of the allocation-free entry point
// (pstm_hip_migrate_tile_noalloc): every GPU the library runs on holds one
// copy for the life of the process.  In host code `scratch` is only a handle
// for hipGetSymbolAddress; its device address, per device, is g_scratch[dev].
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
char* g_scratch[kMaxDevices] = {};

struct DeviceState {
    int device = -1;
    int n_cu = 0;
    void* ws = nullptr;
    size_t ws_size = 0;
};

DeviceState g_state;

// launch shape: time blocks per wave (MW) and traces per LDS batch; 0 = the
// built-in choice.  Set by pstm_hip_set_tuning (the autotuner) or, for
// experiments, the PSTM_HIP_MW / PSTM_HIP_BATCH environment variables.
struct Tuning {
    int mw_cap = 0;
    int batch_cap = 0;
    int pipe = -1;  // pipelined staging: -1 automatic, 0 off, 1 on
};
Tuning g_tune;

struct LastLaunch {
    int mw = 0, batch = 0, grid = 0, items = 0;
};
LastLaunch g_last;

int env_int(const char* name) {
    const char* v = getenv(name);
    return v ? atoi(v) : 0;
}

// optional HIP-event timing of the gather launches (pstm_hip_profile_*)
struct LaunchProfile {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
LaunchProfile g_prof;

int ensure_workspace(size_t bytes) {
    if (bytes <= g_state.ws_size) return 0;
    if (g_state.ws) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(g_state.ws));
        g_state.ws = nullptr;
        g_state.ws_size = 0;
    }
    const size_t sz = std::max(bytes, (size_t)1 << 20);
    HIP_TRY(hipMalloc(&g_state.ws, sz));
    g_state.ws_size = sz;
    return 0;
}

int ensure_init() {
    if (g_state.device >= 0) {
        int cur = -1;
        HIP_TRY(hipGetDevice(&cur));
        if (cur == g_state.device) return 0;
        return fail(4, "pstm_hip: current device differs from the device passed to pstm_hip_init");
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    g_state.device = dev;
    g_state.n_cu = prop.multiProcessorCount;
    return 0;
}

int validate(const PstmMigrateArgs* a) {
    if (!a) return fail(1, "args is NULL");
    if (a->n_traces < 0 || a->n_samples < 2 || a->nx <= 0 || a->ny <= 0 || a->nt <= 0)
        return fail(1, "invalid sizes (need n_traces >= 0, n_samples >= 2, nx, ny, nt > 0)");
    if (!a->x_axis || !a->y_axis || !a->t_axis_ms || !a->vrms || !a->image || !a->fold)
        return fail(1, "axis, velocity, image and fold pointers must be set");
    if (a->n_traces > 0 && (!a->amplitudes || !a->source_x || !a->source_y || !a->receiver_x ||
                            !a->receiver_y || !a->midpoint_x || !a->midpoint_y))
        return fail(1, "trace pointers must be set");
    if (a->vel_mode != PSTM_VEL_1D && a->vel_mode != PSTM_VEL_3D) return fail(1, "invalid vel_mode");
    if (a->traveltime < PSTM_TT_STRAIGHT || a->traveltime > PSTM_TT_CURVED_LEARNED) return fail(1, "invalid traveltime");
    if (a->aa_mode != PSTM_AA_NONE && a->aa_mode != PSTM_AA_TRIANGLE) return fail(1, "invalid aa_mode");
    if (a->aa_mode == PSTM_AA_TRIANGLE && a->traveltime != PSTM_TT_STRAIGHT)
        return fail(1, "triangle anti-aliasing is implemented for the straight-ray traveltime only");
    if (a->aa_mode == PSTM_AA_TRIANGLE && (a->aa_max_half_width < 1 || a->aa_max_half_width > kMaxHalfWidth))
        return fail(1, "aa_max_half_width must be in [1, 64]");
    if (a->traveltime == PSTM_TT_VTI || a->traveltime == PSTM_TT_VTI_LEARNED) {
        if (a->eta_mode < PSTM_ETA_CONST || a->eta_mode > PSTM_ETA_3D) return fail(1, "invalid eta_mode");
        if (a->eta_mode != PSTM_ETA_CONST && !a->eta) return fail(1, "eta pointer must be set");
    }
    if (a->traveltime == PSTM_TT_VTI_LEARNED || a->traveltime == PSTM_TT_CURVED_LEARNED) {
        const int K = a->mlp_node_spacing;
        if (!a->mlp_params) return fail(1, "mlp_params must be set for the learned traveltime");
        if (K < 4 || K > 512) return fail(1, "mlp_node_spacing must be in [4, 512]");
        if (a->mlp_node_growth < 0 || a->mlp_node_growth > 1024)
            return fail(1, "mlp_node_growth must be in [0, 1024]");
        if (a->mlp_node_max != 0 && (a->mlp_node_max < K || a->mlp_node_max > 1024))
            return fail(1, "mlp_node_max must be 0 or in [mlp_node_spacing, 1024]");
        if (a->nt < 2) return fail(1, "the learned traveltime needs at least 2 output samples");
        if (a->mlp_weight_terms < 0 || a->mlp_weight_terms > 2) return fail(1, "mlp_weight_terms must be 0, 1 or 2");
        if (!(a->mlp_tau_max > 0.0) || !(a->mlp_v_hi > a->mlp_v_lo) || !(a->mlp_eta_max > 0.0) ||
            !(a->mlp_out_scale > 0.0))
            return fail(1, "learned traveltime encoding: need tau_max > 0, v_hi > v_lo, eta_max > 0, out_scale > 0");
    }
    if (a->traveltime == PSTM_TT_CURVED_LEARNED && !(fabs(a->curved_k) >= 1e-6))
        return fail(1, "the learned curved-ray traveltime needs a velocity gradient |k| >= 1e-6 (k ~ 0 is the "
                       "straight ray: use PSTM_TT_CURVED)");
    if ((a->traveltime == PSTM_TT_CURVED || a->traveltime == PSTM_TT_CURVED_LEARNED) && !(a->curved_v0 > 0.0))
        return fail(1, "curved_v0 must be positive");
    if (a->interpolation < PSTM_INTERP_LINEAR || a->interpolation > PSTM_INTERP_LANCZOS5)
        return fail(1, "invalid interpolation");
    if (a->interpolation != PSTM_INTERP_LINEAR &&
        (a->traveltime != PSTM_TT_STRAIGHT || a->aa_mode != PSTM_AA_NONE || a->n_effective))
        return fail(1, "interpolation other than linear needs the straight-ray traveltime, no anti-aliasing and "
                       "no effective-contribution counter");
    if (!(a->dt_in_ms > 0.0)) return fail(1, "dt_in_ms must be positive");
    if (!(a->max_aperture_m > 0.0)) return fail(1, "max_aperture_m must be positive");
    if ((long long)a->nx * a->ny * a->nt > (1LL << 40)) return fail(1, "tile too large");
    return 0;
}

// set by plan creation: configure the kernel's dynamic LDS limit only, so that
// nothing but stream work happens while a plan's launch sequence is captured
bool g_attr_only = false;
#endif  // PSTM_GATHER_PART

template <int MW, int TT, bool AA, bool COUNT, bool LOB, bool W2, int INTERP = PSTM_INTERP_LINEAR>
int launch_migrate(const KParams& kp, int grid, size_t lds, hipStream_t st) {
    auto fn = k_migrate<MW, TT, AA, COUNT, LOB, W2, INTERP>;
    static size_t lds_set = 0;  // per instantiation
    if (lds > lds_set) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        lds_set = lds;
    }
    if (pstm_gather_attr_only()) return 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const bool prof = pstm_gather_prof_on();
    if (prof) {
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        HIP_TRY(hipEventRecord(ev0, st));
    }
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), lds, st, kp);
    HIP_TRY(hipGetLastError());
    if (prof) {
        HIP_TRY(hipEventRecord(ev1, st));
        pstm_gather_prof_push(ev0, ev1);
    }
    return 0;
}

template <int TT, bool AA, bool COUNT, bool LOB, bool W2>
int dispatch_mw4(int mw, const KParams& kp, int grid, size_t lds, hipStream_t st) {
    switch (mw) {
        case 1: return launch_migrate<1, TT, AA, COUNT, LOB, W2>(kp, grid, lds, st);
        case 2: return launch_migrate<2, TT, AA, COUNT, LOB, W2>(kp, grid, lds, st);
        case 3: return launch_migrate<3, TT, AA, COUNT, LOB, W2>(kp, grid, lds, st);
        default: return launch_migrate<4, TT, AA, COUNT, LOB, W2>(kp, grid, lds, st);
    }
}

template <int TT, bool AA, bool COUNT, bool LOB>
int dispatch_mw3(int mw, const KParams& kp, int grid, size_t lds, hipStream_t st) {
    return kp.wpow == 2 ? dispatch_mw4<TT, AA, COUNT, LOB, true>(mw, kp, grid, lds, st)
                        : dispatch_mw4<TT, AA, COUNT, LOB, false>(mw, kp, grid, lds, st);
}

template <int TT, bool AA>
int dispatch_mw(int mw, bool count, bool lob, const KParams& kp, int grid, size_t lds, hipStream_t st) {
    if (count) {
        return lob ? dispatch_mw3<TT, AA, true, true>(mw, kp, grid, lds, st)
                   : dispatch_mw3<TT, AA, true, false>(mw, kp, grid, lds, st);
    }
    return lob ? dispatch_mw3<TT, AA, false, true>(mw, kp, grid, lds, st)
               : dispatch_mw3<TT, AA, false, false>(mw, kp, grid, lds, st);
}

#ifndef PSTM_GATHER_PART
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

__global__ void k_reset_queue(unsigned* q) {
    if (threadIdx.x == 0) *q = 0u;
}

// the one-entry tile table of a single-tile call (kernel argument -> device)
__global__ void k_set_view(TileView v, TileView* dst) {
    if (threadIdx.x == 0) *dst = v;
}

// plans: fold += counts taken when the plan was built, one workgroup per tile
__global__ __launch_bounds__(256) void k_fold_add(const TileView* tiles) {
    const TileView t = tiles[blockIdx.x];
    for (int p = threadIdx.x; p < t.nx * t.ny; p += blockDim.x) t.fold[p] += t.fold_cnt[p];
}

struct Shape;
bool lo_check(const PstmMigrateArgs& a, const Shape& sh);

// learned traveltime: the axis lattice (lattice_next) and the largest node
// count of one time window of `win` samples (the kernel's jf .. jl)
std::vector<int> lattice_positions(const PstmMigrateArgs& a) {
    const int kmax = a.mlp_node_max > 0 ? a.mlp_node_max : a.mlp_node_spacing;
    std::vector<int> pos{0};
    while (pos.back() < a.nt - 1)
        pos.push_back(lattice_next(pos.back(), a.mlp_node_spacing, a.mlp_node_growth, kmax, a.nt - 1));
    return pos;
}

int lattice_window_nodes(const std::vector<int>& pos, int nt, int win) {
    const int total = (int)pos.size();
    auto upper = [&](int k) { return (int)(std::upper_bound(pos.begin(), pos.end(), k) - pos.begin()); };
    int most = 0;
    for (int kb = 0; kb < nt; kb += win) {
        const int kend = std::min(kb + win, nt);
        const int jf = std::min(upper(kb) - 1, total - 2);
        const int jl = std::max(upper(kend - 2), jf + 1);
        most = std::max(most, jl - jf + 1);
    }
    return most;
}

// Launch shape and LDS layout shared by every tile of a call or plan.
struct Shape {
    int tt = 0;                  // traveltime evaluated (curved with |k| < 1e-10 -> straight)
    bool aa = false, learned = false, vti_like = false;
    int lmax = 0, lo = 0, rowlen = 0, pair_vec = 1;
    size_t elem = 4;
    int mw = 1, batch = 1, lat_nodes = 0, win = 256, nwin = 1;
    int lat_total = 0;           // learned: lattice nodes over the output axis
    int mlp_terms = 1;           // learned: bf16 terms per weight
    size_t tot_bytes = 0, lds = 0;
    int occupancy = 1;
    bool pipe = false;
    int interp = PSTM_INTERP_LINEAR;
};

Shape launch_shape(const PstmMigrateArgs& a) {
    Shape sh;
    sh.tt = a.traveltime;
    if (sh.tt == PSTM_TT_CURVED && fabs(a.curved_k) < 1e-10) sh.tt = PSTM_TT_STRAIGHT;
    sh.aa = a.aa_mode == PSTM_AA_TRIANGLE;
    sh.learned = learned_tt(sh.tt);
    sh.vti_like = sh.tt == PSTM_TT_VTI || sh.learned;
    sh.lmax = sh.aa ? a.aa_max_half_width : 0;
    sh.lo = sh.aa ? sh.lmax + 1 : 0;
    const int ns = a.n_samples, nt = a.nt;
    sh.rowlen = sh.aa ? ns + 2 * sh.lmax + 2 : ns;
    sh.elem = sh.aa ? sizeof(double) : sizeof(float);
    sh.pair_vec = (sh.aa || sh.tt == PSTM_TT_CURVED) ? 2 : 1;
    // time windows of at most 1024 samples, split evenly; MW blocks of 64 per wave
    const int nwin = (nt + 1023) / 1024;
    const int per_win = (nt + nwin - 1) / nwin;
    int mw_cap = g_tune.mw_cap > 0 ? g_tune.mw_cap : env_int("PSTM_HIP_MW");
    if (mw_cap <= 0) mw_cap = 4;  // measured best on the reflector job, with and without AA
    mw_cap = std::min(4, mw_cap);
    sh.mw = std::min(mw_cap, std::max(1, (per_win + 255) / 256));
    sh.interp = a.interpolation;
    if (sh.interp != PSTM_INTERP_LINEAR) sh.mw = interp_mw(sh.interp);  // their one instantiation
    sh.mlp_terms = sh.learned && a.mlp_weight_terms == 2 ? 2 : 1;
    // pipelined staging: f32 rows through kPF registers per thread (the batch's
    // rows must fit), AA rows by direct global -> LDS loads into a second row
    // buffer (LDS holds two buffers, so the batch shrinks); automatic: off
    int pipe = g_tune.pipe;
    if (pipe < 0 && getenv("PSTM_HIP_PIPE")) pipe = atoi(getenv("PSTM_HIP_PIPE")) != 0;
    // off unless asked for: direct loads at the batch start let the batch
    // grow to what the LDS holds (f32 rows through registers cap it at kPF
    // dwords per thread: 6 traces of 2001 samples) and measured faster
    if (pipe < 0) pipe = 0;
    const int rl_dw = sh.rowlen * (int)(sh.elem / 4);
    const int pipe_batch = sh.aa ? kMaxBatch : kPF * kThreads / rl_dw;
    // (the learned traveltime always stages by direct loads at the batch start)
    sh.pipe = pipe == 1 && pipe_batch >= 1 && !sh.learned;
    const int nbuf = (sh.pipe && sh.aa) ? 2 : 1;
    // LDS: trace rows + pair table + staged pair inputs + f64 totals; learned: + lattice, widest
    // apertures, row list, node inputs, network operands (kept for the
    // workgroup's life)
    // learned: per (evaluated pillar, trace, extended node) the network
    // outputs (8 B) and the row list entry (2 B), per (pillar, trace, segment) the
    // Hermite coefficients of both legs (32 B), per (pillar, extended node)
    // the widest aperture (4 B) and the network inputs (16 B)
    auto lds_for = [&](int b) {
        const size_t nx = (size_t)sh.lat_nodes + 2;
        return align_up((size_t)nbuf * align_up((size_t)b * sh.rowlen * sh.elem, 16) +
                            (size_t)kNPil * b * sh.pair_vec * 16, 16) +
               kStageBytes + sh.tot_bytes +
               (sh.learned ? align_up((size_t)3 * b * nx * 8, 16) + (size_t)kNPil * b * (sh.lat_nodes - 1) * 32 +
                                 align_up((size_t)kNPil * nx * 4, 16) + align_up((size_t)3 * b * nx * 2, 16) +
                                 (size_t)kNPil * nx * 16 + (size_t)sh.mlp_terms * kAVec8 * 16 +
                                 (size_t)kBVals * 4 + align_up((size_t)sh.lat_total * 4, 16)
                           : 0);
    };
    std::vector<int> npos;
    if (sh.learned) {
        npos = lattice_positions(a);
        sh.lat_total = (int)npos.size();
    }
    // sample blocks per wave: the learned layout must fit one trace per batch
    for (;;) {
        // running totals: f64 in LDS (the learned traveltime keeps f32 totals in registers)
        sh.tot_bytes = sh.learned ? 0 : (size_t)kPW * sh.mw * kThreads * sizeof(double);
        sh.lat_nodes = sh.learned ? lattice_window_nodes(npos, nt, 256 * sh.mw) : 0;
        if (sh.mw == 1 || lds_for(1) <= kLdsBudget) break;
        --sh.mw;
    }
    int batch_cap = g_tune.batch_cap > 0 ? g_tune.batch_cap : env_int("PSTM_HIP_BATCH");
    sh.batch = batch_cap > 0 ? std::min(batch_cap, kMaxBatch) : kMaxBatch;
    if (sh.pipe) sh.batch = std::min(sh.batch, pipe_batch);
    while (sh.batch > 1 && lds_for(sh.batch) > kLdsBudget) --sh.batch;
    sh.lds = lds_for(sh.batch);
    sh.win = 256 * sh.mw;
    sh.nwin = (nt + sh.win - 1) / sh.win;
    sh.occupancy = sh.lds <= kLdsBudget / 2 ? 2 : 1;  // workgroups per CU the LDS allows
    return sh;
}

bool lo_check(const PstmMigrateArgs& a, const Shape& sh) {
    // (the learned times are never below t0 >= 0, nor below 0: u >= 0 unless the input starts late)
    return a.t_start_in_ms > 0.0 || (sh.tt == PSTM_TT_CURVED && a.curved_k < 0.0) || (sh.vti_like && !sh.learned);
}

// Kernel parameters common to every tile (everything but the tile table).
void fill_common(KParams& kp, const PstmMigrateArgs& a, const Shape& sh, double tan_dip) {
    const int ns = a.n_samples, nt = a.nt;
    kp.tiles = nullptr;
    kp.n_tiles = 0;
    kp.amp = a.amplitudes;
    kp.prow_global = 0;
    kp.ns = ns;
    kp.nt = nt;
    kp.rowlen = sh.rowlen;
    kp.lo = sh.lo;
    kp.batch = sh.batch;
    kp.pipe = sh.pipe ? 1 : 0;
    kp.rows_bytes = (int)align_up((size_t)sh.batch * sh.rowlen * sh.elem, 16);
    kp.pair_off = kp.rows_bytes * ((sh.pipe && sh.aa) ? 2 : 1);
    kp.stage_off = (int)align_up((size_t)kp.pair_off + (size_t)kNPil * sh.batch * sh.pair_vec * 16, 16);
    kp.tot_off = kp.stage_off + (int)kStageBytes;
    kp.taper_off = a.taper_fraction > 0.0 ? (float)(-(1.0 - a.taper_fraction) / a.taper_fraction) : -1e30f;
    kp.wpow = (a.apply_spreading ? 1 : 0) + (a.apply_obliquity ? 1 : 0);
    kp.interp = a.interpolation;
    const double dt_s = a.dt_in_ms / 1000.0;
    kp.idt = (float)(1.0 / dt_s);
    kp.uoff = (float)(-(a.t_start_in_ms / 1000.0) / dt_s);
    kp.umax = (float)(ns - 1);
    kp.max_ap = (float)a.max_aperture_m;
    kp.gdx_idt = (float)(a.grid_dx / dt_s);
    kp.gdy_idt = (float)(a.grid_dy / dt_s);
    kp.lmax = (float)sh.lmax;
    kp.ck2 = (float)(a.curved_k * a.curved_k);
    kp.ik_ln2 = (float)(sh.tt == PSTM_TT_CURVED ? M_LN2 / a.curved_k : 0.0);
    kp.v0 = (float)a.curved_v0;
    kp.t_floor = a.curved_k > 0.0 ? 0.0f : -INFINITY;
    kp.sx64 = a.source_x;
    kp.sy64 = a.source_y;
    kp.rx64 = a.receiver_x;
    kp.ry64 = a.receiver_y;
    kp.mx64 = a.midpoint_x;
    kp.my64 = a.midpoint_y;
    kp.tax64 = a.t_axis_ms;
    kp.tt_req = a.traveltime;
    // u >= 0 can only fail or be decided differently from f64 when t can be
    // at or below t_start: t_start > 0, a curved ray with k < 0, or VTI
    // (t = t0 for t0 <= 0, and negative output times give u < 0)
    kp.eps_u = (float)(2e-3 + 1e-5 * (double)(ns - 1));
    // the straight-ray f32 position is within ~2.6e-7 u of the f64 one (CPU
    // emulation of this arithmetic at 800 m tiles, 2500 m apertures, 3 km
    // offsets); the fraction-0.5 band keeps a 4x margin
    kp.eps_half = 1e-6f;
    kp.lo_eps = lo_check(a, sh) ? kp.eps_u : -INFINITY;
    kp.umax_lo = kp.umax - kp.eps_u;
    kp.umax_hi = kp.umax + kp.eps_u;
    kp.eps_x = a.n_effective ? 1e-4f : 0.0f;
    kp.tan_dip = tan_dip;
    kp.min_ap64 = a.min_aperture_m;
    kp.max_ap64 = a.max_aperture_m;
    kp.t_start_s = a.t_start_in_ms / 1000.0;
    kp.dt_s = dt_s;
    kp.cv0 = a.curved_v0;
    kp.ck = a.curved_k;
    kp.eta_const = a.eta_constant;
    const int nx = sh.lat_nodes + 2;  // learned: extended nodes of a window
    kp.lat_off = (int)(kp.tot_off + sh.tot_bytes);
    kp.lat_nodes = sh.lat_nodes;
    kp.coef_off = kp.lat_off + (int)align_up((size_t)3 * sh.batch * nx * 8, 16);
    kp.lat_cimin_off = kp.coef_off + kNPil * sh.batch * (sh.lat_nodes - 1) * 32;
    kp.lat_list_off = kp.lat_cimin_off + (int)align_up((size_t)kNPil * nx * 4, 16);
    kp.node_off = kp.lat_list_off + (int)align_up((size_t)3 * sh.batch * nx * 2, 16);
    kp.mlp_off = kp.node_off + kNPil * nx * 16;
    kp.lat_pos_off = kp.mlp_off + (int)((size_t)sh.mlp_terms * kAVec8 * 16 + (size_t)kBVals * 4);
    kp.lat_total = sh.lat_total;
    kp.lat_kmin = a.mlp_node_spacing;
    kp.lat_grow = a.mlp_node_growth;
    kp.lat_kmax = a.mlp_node_max > 0 ? a.mlp_node_max : a.mlp_node_spacing;
    kp.m_tau_sc = sh.learned ? (float)(2.0 / a.mlp_tau_max) : 0.0f;
    kp.m_v_lo = (float)a.mlp_v_lo;
    kp.m_v_sc = sh.learned ? (float)(2.0 / (a.mlp_v_hi - a.mlp_v_lo)) : 0.0f;
    kp.m_eta_sc = sh.learned ? (float)(2.0 / a.mlp_eta_max) : 0.0f;
    kp.m_inv_out = sh.learned ? (float)(1.0 / a.mlp_out_scale) : 0.0f;
    kp.mlpA = nullptr;
    kp.mlpB = nullptr;
    kp.mlp_split = sh.learned && a.mlp_weight_terms == 2;
}

// Per-tile buffers: f32 tile-relative geometry (ntr x 32 B), per-(pillar,
// sample) terms (npv x nt x 16 B) and VTI eta terms (nev x nt x 8 B).
struct TileBufs {
    float4* geo;
    float4* kc;
    float2* ek;
};

size_t tile_bytes(const PstmMigrateArgs& a, const Shape& sh, int ntr) {
    const int npil = a.nx * a.ny;
    const int npv = a.vel_mode == PSTM_VEL_3D ? npil : 1;
    const int nev = a.eta_mode == PSTM_ETA_3D ? npil : 1;
    return align_up((size_t)ntr * 32, 256) + align_up((size_t)npv * a.nt * 16, 256) +
           align_up(sh.vti_like ? (size_t)nev * a.nt * 8 : 0, 256);
}

TileBufs carve_tile(char* base, const PstmMigrateArgs& a, const Shape& sh, int ntr) {
    const int npil = a.nx * a.ny;
    const int npv = a.vel_mode == PSTM_VEL_3D ? npil : 1;
    TileBufs b;
    b.geo = reinterpret_cast<float4*>(base);
    base += align_up((size_t)ntr * 32, 256);
    b.kc = reinterpret_cast<float4*>(base);
    base += align_up((size_t)npv * a.nt * 16, 256);
    b.ek = reinterpret_cast<float2*>(base);
    return b;
}

// Launch the per-tile preparation kernels of `a` (one tile: axes, velocity,
// eta, trace list) into b; qk/t0k (shared by all tiles) when non-null.
int prep_tile(const PstmMigrateArgs& a, const Shape& sh, double tan_dip, const TileBufs& b, float* qk,
              float* t0k, hipStream_t st) {
    const int npil = a.nx * a.ny;
    const int npv = a.vel_mode == PSTM_VEL_3D ? npil : 1;
    const int nev = a.eta_mode == PSTM_ETA_3D ? npil : 1;
    if (a.n_traces > 0) {
        hipLaunchKernelGGL(k_prep_geometry, dim3((a.n_traces + 255) / 256), dim3(256), 0, st, a, b.geo);
        HIP_TRY(hipGetLastError());
    }
    const long nkc = (long)npv * a.nt;
    hipLaunchKernelGGL(k_prep_time_terms, dim3((unsigned)((nkc + 255) / 256)), dim3(256), 0, st, a, sh.tt, npv,
                       tan_dip, b.kc, qk, t0k);
    HIP_TRY(hipGetLastError());
    if (sh.vti_like) {
        const long nek = (long)nev * a.nt;
        hipLaunchKernelGGL(k_prep_eta, dim3((unsigned)((nek + 255) / 256)), dim3(256), 0, st, a, nev, b.ek);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

TileView make_view(const PstmMigrateArgs& a, const Shape& sh, const TileBufs& b, int item0) {
    const int npil = a.nx * a.ny;
    TileView t;
    t.xax = a.x_axis;
    t.yax = a.y_axis;
    t.kc = b.kc;
    t.ek = b.ek;
    t.vrms64 = a.vrms;
    t.eta64 = (sh.vti_like && a.eta_mode != PSTM_ETA_CONST) ? a.eta : nullptr;
    t.image = a.image;
    t.geo = b.geo;
    t.tidx = a.trace_index;
    t.fold = a.fold;
    t.fold_cnt = nullptr;
    t.ntr = a.n_traces;
    t.nx = a.nx;
    t.ny = a.ny;
    t.kc_pstride = a.vel_mode == PSTM_VEL_3D ? a.nt : 0;
    t.ek_pstride = (sh.vti_like && a.eta_mode == PSTM_ETA_3D) ? a.nt : 0;
    t.eta_pstride = a.eta_mode == PSTM_ETA_3D ? a.nt : 0;
    t.n_pgroups = (npil + kNPil - 1) / kNPil;
    t.item0 = item0;
    return t;
}

long tile_items(const PstmMigrateArgs& a, const Shape& sh) {
    return (long)((a.nx * a.ny + kNPil - 1) / kNPil) * sh.nwin;
}

#endif  // PSTM_GATHER_PART

// straight ray, no anti-aliasing, no counter, MW = kInterpMW (launch_shape),
// every lower-bound check: the one instantiation per interpolator
template <int INTERP>
int launch_interp(const KParams& kp, int grid, size_t lds, hipStream_t st) {
    return launch_migrate<kInterpMW<INTERP>, PSTM_TT_STRAIGHT, false, false, true, false, INTERP>(kp, grid, lds,
                                                                                                  st);
}

#ifndef PSTM_GATHER_PART
// the gather launch of a call or plan, through the launcher of its kernel group
int dispatch(const Shape& sh, bool count, bool lob, const KParams& kp, int grid, hipStream_t st) {
    int (*part)(int, bool, bool, int, const void*, int, size_t, hipStream_t);
    if (sh.interp != PSTM_INTERP_LINEAR) part = pstm_gather_part6;
    else if (sh.tt == PSTM_TT_STRAIGHT) part = sh.aa ? pstm_gather_part1 : pstm_gather_part2;
    else if (sh.tt == PSTM_TT_CURVED) part = pstm_gather_part3;
    else if (sh.tt == PSTM_TT_VTI) part = pstm_gather_part4;
    else if (sh.tt == PSTM_TT_CURVED_LEARNED) part = pstm_gather_part7;
    else part = pstm_gather_part5;
    return part(sh.mw, count, lob, sh.interp, &kp, grid, sh.lds, st);
}

// shared by migrate_tile and migrate_bins: `ntr` rows starting at row
// offset `row0` of args->trace_index (or of the block).  The temporaries
// live in the library's workspace (grown with hipMalloc when too small), or,
// when fixed_ws is set, in that buffer of fixed_ws_size bytes and nowhere
// else (status 1 when they do not fit).
int run_migration(const PstmMigrateArgs* in, int row0, int ntr, int do_fold, hipStream_t st,
                  char* fixed_ws = nullptr, size_t fixed_ws_size = 0) {
    int rc = validate(in);
    if (rc) return rc;
    rc = ensure_init();
    if (rc) return rc;
    PstmMigrateArgs a = *in;
    a.n_traces = ntr;
    if (a.trace_index) a.trace_index += row0;
    else if (row0 > 0) {
        // rows [row0, row0+ntr) of the block: shift the per-trace arrays
        a.amplitudes += (size_t)row0 * a.n_samples;
        a.source_x += row0; a.source_y += row0;
        a.receiver_x += row0; a.receiver_y += row0;
        a.midpoint_x += row0; a.midpoint_y += row0;
        if (a.trace_weights) a.trace_weights += row0;
    }
    if (ntr == 0) return 0;
    const int nt = a.nt;
    const Shape sh = launch_shape(a);
    if (sh.lds > kLdsBudget) return fail(5, "trace too long for one LDS row (n_samples too large)");
    const long n_items = tile_items(a, sh);
    if (n_items > (1L << 30)) return fail(1, "tile too large");
    const int grid = (int)std::min<long>(n_items, (long)g_state.n_cu * sh.occupancy);
    g_last.mw = sh.mw;
    g_last.batch = sh.batch;
    g_last.grid = grid;
    g_last.items = (int)n_items;

    // workspace layout
    size_t off = 0;
    const size_t o_tile = off; off = align_up(off + tile_bytes(a, sh, ntr), 256);
    const size_t o_qk = off; off = align_up(off + (size_t)nt * 4, 256);
    const size_t o_t0 = off; off = align_up(off + (size_t)nt * 4, 256);
    const size_t o_mlpA = off; off = align_up(off + (sh.learned ? (size_t)kAVals * 4 : 0), 256);
    const size_t o_mlpB = off; off = align_up(off + (sh.learned ? (size_t)kBVals * 4 : 0), 256);
    const size_t o_cnt = off; off = align_up(off + (size_t)grid * 8, 256);
    const size_t o_q = off; off = align_up(off + 8, 256);
    const size_t o_view = off; off = align_up(off + sizeof(TileView), 256);
    const size_t o_prow = off; off = align_up(off + (sh.aa ? (size_t)ntr * sh.rowlen * 8 : 0), 256);
    char* ws = fixed_ws;
    if (fixed_ws) {
        if (off > fixed_ws_size)
            return fail(1, "this migration needs " + std::to_string(off) + " bytes of temporaries; the "
                           "allocation-free entry point holds at most " + std::to_string(fixed_ws_size) +
                           " (fewer traces or pillars per call, or 1-D velocity)");
    } else {
        rc = ensure_workspace(off);
        if (rc) return rc;
        ws = static_cast<char*>(g_state.ws);
    }
    const TileBufs tb = carve_tile(ws + o_tile, a, sh, ntr);
    float* qk = reinterpret_cast<float*>(ws + o_qk);
    float* t0k = reinterpret_cast<float*>(ws + o_t0);
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(ws + o_cnt);
    unsigned* queue = reinterpret_cast<unsigned*>(ws + o_q);
    double* prow = reinterpret_cast<double*>(ws + o_prow);

    const double tan_dip = tan(a.max_dip_deg * M_PI / 180.0);
    rc = prep_tile(a, sh, tan_dip, tb, qk, t0k, st);
    if (rc) return rc;
    if (sh.learned) {
        const int n = kAVals + kBVals;
        hipLaunchKernelGGL(k_prep_mlp, dim3((n + 255) / 256), dim3(256), 0, st, a.mlp_params,
                           reinterpret_cast<unsigned short*>(ws + o_mlpA), reinterpret_cast<float*>(ws + o_mlpB));
        HIP_TRY(hipGetLastError());
    }
    if (sh.aa) {
        hipLaunchKernelGGL(k_integrate, dim3(ntr), dim3(256), 0, st, a, sh.rowlen, sh.lo, prow);
        HIP_TRY(hipGetLastError());
    }
    if (do_fold) {
        hipLaunchKernelGGL(k_fold, dim3(a.nx * a.ny), dim3(256), 0, st, a);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, st, queue);
    HIP_TRY(hipGetLastError());
    TileView* view = reinterpret_cast<TileView*>(ws + o_view);
    hipLaunchKernelGGL(k_set_view, dim3(1), dim3(64), 0, st, make_view(a, sh, tb, 0), view);
    HIP_TRY(hipGetLastError());

    KParams kp;
    fill_common(kp, a, sh, tan_dip);
    kp.tiles = view;
    kp.n_tiles = 1;
    kp.queue = queue;
    kp.prow = prow;
    kp.qk = qk;
    kp.t0k = t0k;
    kp.cnt_part = a.n_effective ? cnt : nullptr;
    kp.n_items = (int)n_items;
    if (sh.learned) {
        kp.mlpA = ws + o_mlpA;
        kp.mlpB = ws + o_mlpB;
    }
    const bool count = a.n_effective != nullptr;
    rc = dispatch(sh, count, kp.lo_eps != -INFINITY, kp, grid, st);
    if (rc) return rc;
    if (a.n_effective) {
        hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(256), 0, st, (const unsigned long long*)cnt, grid,
                           a.n_effective);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

}  // namespace

// Failure message setter for the library's other translation units (the
// output kernels), so that pstm_hip_last_error() reports their failures too.
__attribute__((visibility("hidden"))) void pstm_set_last_error(const std::string& msg) { g_last_error = msg; }
__attribute__((visibility("hidden"))) bool pstm_gather_attr_only() { return g_attr_only; }
__attribute__((visibility("hidden"))) bool pstm_gather_prof_on() { return g_prof.on; }
__attribute__((visibility("hidden"))) void pstm_gather_prof_push(hipEvent_t ev0, hipEvent_t ev1) {
    g_prof.events.emplace_back(ev0, ev1);
}

// A device-resident work queue over many output tiles (common-offset bins
// included: a tile's trace list may hold rows of every bin): per-tile
// geometry, velocity and fold terms are built once; a run is one persistent
// gather launch whose workgroups take (tile, pillar group, time window)
// items from a device counter, heaviest tiles first, plus a few small
// kernels, replayed from a captured hipGraph.
struct PstmPlan {
    PstmMigrateArgs common;
    Shape sh;
    int n_tiles = 0;
    long n_items = 0;
    int grid = 0;
    int block_rows = 0;
    double tan_dip = 0.0;
    void* mem = nullptr;
    TileView* d_views = nullptr;
    unsigned* queue = nullptr;
    unsigned long long* cnt = nullptr;
    double* prow = nullptr;
    void* mlpA = nullptr;
    void* mlpB = nullptr;
    KParams kp;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// the launch sequence of one plan run on `st`
int plan_enqueue(PstmPlan* pl, hipStream_t st) {
    const PstmMigrateArgs& a = pl->common;
    const Shape& sh = pl->sh;
    if (sh.learned) {
        const int n = kAVals + kBVals;
        hipLaunchKernelGGL(k_prep_mlp, dim3((n + 255) / 256), dim3(256), 0, st, a.mlp_params,
                           static_cast<unsigned short*>(pl->mlpA), static_cast<float*>(pl->mlpB));
        HIP_TRY(hipGetLastError());
    }
    if (sh.aa) {
        PstmMigrateArgs all = a;
        all.trace_index = nullptr;
        all.n_traces = pl->block_rows;
        hipLaunchKernelGGL(k_integrate, dim3(pl->block_rows), dim3(256), 0, st, all, sh.rowlen, sh.lo, pl->prow);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_fold_add, dim3(pl->n_tiles), dim3(256), 0, st, (const TileView*)pl->d_views);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, st, pl->queue);
    HIP_TRY(hipGetLastError());
    const bool count = a.n_effective != nullptr;
    int rc = dispatch(sh, count, pl->kp.lo_eps != -INFINITY, pl->kp, pl->grid, st);
    if (rc) return rc;
    if (count) {
        hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(256), 0, st, (const unsigned long long*)pl->cnt, pl->grid,
                           a.n_effective);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

void plan_free(PstmPlan* pl) {
    if (!pl) return;
    // A run may still be in flight on the caller's stream (HipPlan.run returns
    // at once): wait before destroying its executable graph, which holds the
    // gather's kernel arguments the running kernel still reads, and before
    // freeing the plan's memory.
    (void)hipDeviceSynchronize();
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->cap) (void)hipStreamDestroy(pl->cap);
    if (pl->mem) (void)hipFree(pl->mem);
    delete pl;
}

// Byte ranges [image, image + nx ny nt f32) and [fold, fold + nx ny i32) of
// the tiles of a plan must not overlap one another (an image range may not
// overlap a fold range either): sort the ranges by start; a range that starts
// before the furthest end seen so far overlaps an earlier one.
int check_disjoint(const std::vector<PstmMigrateArgs>& targs) {
    struct Range { uintptr_t lo, hi; int tile; const char* what; };
    std::vector<Range> r;
    r.reserve(targs.size() * 2);
    for (size_t i = 0; i < targs.size(); ++i) {
        const PstmMigrateArgs& t = targs[i];
        const uintptr_t im = reinterpret_cast<uintptr_t>(t.image), fo = reinterpret_cast<uintptr_t>(t.fold);
        r.push_back({im, im + (uintptr_t)t.nx * t.ny * t.nt * sizeof(float), (int)i, "image"});
        r.push_back({fo, fo + (uintptr_t)t.nx * t.ny * sizeof(int32_t), (int)i, "fold"});
    }
    std::sort(r.begin(), r.end(), [](const Range& a, const Range& b) { return a.lo < b.lo; });
    size_t far = 0;  // range with the furthest end so far
    for (size_t k = 1; k < r.size(); ++k) {
        if (r[k].lo < r[far].hi)
            return fail(1, std::string("pstm_hip_plan_create: the ") + r[far].what + " of tile " +
                               std::to_string(r[far].tile) + " overlaps the " + r[k].what + " of tile " +
                               std::to_string(r[k].tile) + " (every tile needs its own image and fold memory)");
        if (r[k].hi > r[far].hi) far = k;
    }
    return 0;
}

}  // namespace

extern "C" {

int pstm_hip_plan_create(const PstmMigrateArgs* common, const PstmTileDesc* tiles, int32_t n_tiles,
                         PstmPlan** out) {
    g_last_error.clear();
    if (!common || !tiles || !out || n_tiles <= 0) return fail(1, "pstm_hip_plan_create: bad arguments");
    *out = nullptr;
    // per-tile argument sets, validated like single-tile calls (host checks
    // only, before the device is touched)
    int rc = 0;
    std::vector<PstmMigrateArgs> targs(n_tiles);
    for (int i = 0; i < n_tiles; ++i) {
        PstmMigrateArgs t = *common;
        const PstmTileDesc& d = tiles[i];
        t.x_axis = d.x_axis;
        t.y_axis = d.y_axis;
        t.vrms = d.vrms;
        t.eta = d.eta;
        t.image = d.image;
        t.fold = d.fold;
        t.trace_index = d.trace_index;
        t.n_traces = d.n_traces;
        t.nx = d.nx;
        t.ny = d.ny;
        rc = validate(&t);
        if (rc) return fail(rc, "tile " + std::to_string(i) + ": " + g_last_error);
        if (d.n_traces > 0 && !d.trace_index) return fail(1, "tile " + std::to_string(i) + ": trace_index must be set");
        targs[i] = t;
    }
    if (common->n_traces <= 0) return fail(1, "pstm_hip_plan_create: common.n_traces must be the rows of the block");
    // the gather adds into image cells and k_fold_add into fold cells without
    // atomics: each tile must own its image and fold memory
    rc = check_disjoint(targs);
    if (rc) return rc;
    rc = ensure_init();
    if (rc) return rc;
    const Shape sh = launch_shape(targs[0]);
    if (sh.lds > kLdsBudget) return fail(5, "trace too long for one LDS row (n_samples too large)");
    // heaviest tiles first
    std::vector<int> order(n_tiles);
    for (int i = 0; i < n_tiles; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return (long)targs[x].n_traces * tile_items(targs[x], sh) >
                                                (long)targs[y].n_traces * tile_items(targs[y], sh); });
    auto* pl = new PstmPlan();
    pl->common = *common;
    pl->sh = sh;
    pl->n_tiles = n_tiles;
    pl->block_rows = common->n_traces;
    pl->tan_dip = tan(common->max_dip_deg * M_PI / 180.0);
    // device memory: tiles, fold counts, shared terms, views, counters, rows
    std::vector<size_t> o_tile(n_tiles), o_fc(n_tiles);
    size_t off = 0;
    for (int i = 0; i < n_tiles; ++i) {
        o_tile[i] = off;
        off = align_up(off + tile_bytes(targs[i], sh, targs[i].n_traces), 256);
        o_fc[i] = off;
        off = align_up(off + (size_t)targs[i].nx * targs[i].ny * 4, 256);
    }
    const int nt = common->nt;
    const size_t o_qk = off; off = align_up(off + (size_t)nt * 4, 256);
    const size_t o_t0 = off; off = align_up(off + (size_t)nt * 4, 256);
    const size_t o_views = off; off = align_up(off + (size_t)n_tiles * sizeof(TileView), 256);
    long n_items = 0;
    for (int i = 0; i < n_tiles; ++i) n_items += tile_items(targs[i], sh);
    if (n_items > (1L << 30)) {
        plan_free(pl);
        return fail(1, "plan too large");
    }
    pl->n_items = n_items;
    pl->grid = (int)std::min<long>(n_items, (long)g_state.n_cu * sh.occupancy);
    const size_t o_cnt = off; off = align_up(off + (size_t)pl->grid * 8, 256);
    const size_t o_q = off; off = align_up(off + 8, 256);
    const size_t o_mlpA = off; off = align_up(off + (sh.learned ? (size_t)kAVals * 4 : 0), 256);
    const size_t o_mlpB = off; off = align_up(off + (sh.learned ? (size_t)kBVals * 4 : 0), 256);
    const size_t o_prow = off; off = align_up(off + (sh.aa ? (size_t)pl->block_rows * sh.rowlen * 8 : 0), 256);
    hipError_t e = hipMalloc(&pl->mem, off);
    if (e != hipSuccess) {
        pl->mem = nullptr;
        plan_free(pl);
        return fail(3, std::string("pstm_hip_plan_create: hipMalloc: ") + hipGetErrorString(e));
    }
    char* base = static_cast<char*>(pl->mem);
    float* qk = reinterpret_cast<float*>(base + o_qk);
    float* t0k = reinterpret_cast<float*>(base + o_t0);
    pl->d_views = reinterpret_cast<TileView*>(base + o_views);
    pl->cnt = reinterpret_cast<unsigned long long*>(base + o_cnt);
    pl->queue = reinterpret_cast<unsigned*>(base + o_q);
    pl->mlpA = sh.learned ? base + o_mlpA : nullptr;
    pl->mlpB = sh.learned ? base + o_mlpB : nullptr;
    pl->prow = reinterpret_cast<double*>(base + o_prow);
    e = hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        pl->cap = nullptr;
        plan_free(pl);
        return fail(3, std::string("pstm_hip_plan_create: hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipStream_t st = pl->cap;
    // per-tile preparation, fold counts and views (in item order)
    std::vector<TileView> views(n_tiles);
    long item0 = 0;
    for (int r = 0; r < n_tiles; ++r) {
        const int i = order[r];
        const PstmMigrateArgs& t = targs[i];
        const TileBufs tb = carve_tile(base + o_tile[i], t, sh, t.n_traces);
        rc = prep_tile(t, sh, pl->tan_dip, tb, r == 0 ? qk : nullptr, r == 0 ? t0k : nullptr, st);
        if (rc) { plan_free(pl); return rc; }
        int32_t* fc = reinterpret_cast<int32_t*>(base + o_fc[i]);
        e = hipMemsetAsync(fc, 0, (size_t)t.nx * t.ny * 4, st);
        if (e != hipSuccess) { plan_free(pl); return fail(3, hipGetErrorString(e)); }
        if (t.n_traces > 0) {
            PstmMigrateArgs fa = t;
            fa.fold = fc;
            hipLaunchKernelGGL(k_fold, dim3(t.nx * t.ny), dim3(256), 0, st, fa);
            e = hipGetLastError();
            if (e != hipSuccess) { plan_free(pl); return fail(3, hipGetErrorString(e)); }
        }
        views[r] = make_view(t, sh, tb, (int)item0);
        views[r].fold_cnt = fc;
        item0 += tile_items(t, sh);
    }
    e = hipMemcpyAsync(pl->d_views, views.data(), views.size() * sizeof(TileView), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { plan_free(pl); return fail(3, std::string("pstm_hip_plan_create: ") + hipGetErrorString(e)); }

    KParams& kp = pl->kp;
    fill_common(kp, *common, sh, pl->tan_dip);
    kp.tiles = pl->d_views;
    kp.n_tiles = n_tiles;
    kp.queue = pl->queue;
    kp.prow = pl->prow;
    kp.prow_global = 1;
    kp.qk = qk;
    kp.t0k = t0k;
    kp.cnt_part = common->n_effective ? pl->cnt : nullptr;
    kp.n_items = (int)n_items;
    kp.mlpA = pl->mlpA;
    kp.mlpB = pl->mlpB;
    // kernel attributes now, so that a captured run holds stream work only
    g_attr_only = true;
    rc = dispatch(sh, common->n_effective != nullptr, kp.lo_eps != -INFINITY, kp, pl->grid, st);
    g_attr_only = false;
    if (rc) { plan_free(pl); return rc; }
    g_last.mw = sh.mw;
    g_last.batch = sh.batch;
    g_last.grid = pl->grid;
    g_last.items = (int)n_items;
    *out = pl;
    return 0;
}

int pstm_hip_plan_run(PstmPlan* pl, int32_t use_graph, void* stream) {
    g_last_error.clear();
    if (!pl) return fail(1, "pstm_hip_plan_run: plan is NULL");
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!use_graph) return plan_enqueue(pl, st);
    if (!pl->exec) {
        const bool prof = g_prof.on;
        g_prof.on = false;  // no event records inside the captured sequence
        HIP_TRY(hipStreamBeginCapture(pl->cap, hipStreamCaptureModeThreadLocal));
        rc = plan_enqueue(pl, pl->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(pl->cap, &g);
        g_prof.on = prof;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIP_TRY(e);
        pl->graph = g;
        HIP_TRY(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
    }
    if (g_prof.on) {
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        HIP_TRY(hipEventRecord(ev0, st));
    }
    HIP_TRY(hipGraphLaunch(pl->exec, st));
    if (g_prof.on) {
        HIP_TRY(hipEventRecord(ev1, st));
        g_prof.events.emplace_back(ev0, ev1);
    }
    return 0;
}

int pstm_hip_plan_info(const PstmPlan* pl, long long* n_items, int* grid, int* mw, int* batch) {
    g_last_error.clear();
    if (!pl || !n_items || !grid || !mw || !batch) return fail(1, "pstm_hip_plan_info: bad arguments");
    *n_items = pl->n_items;
    *grid = pl->grid;
    *mw = pl->sh.mw;
    *batch = pl->sh.batch;
    return 0;
}

int pstm_hip_plan_destroy(PstmPlan* pl) {
    g_last_error.clear();
    plan_free(pl);
    return 0;
}


int pstm_hip_abi_version(void) { return PSTM_HIP_ABI_VERSION; }

size_t pstm_hip_args_size(void) { return sizeof(PstmMigrateArgs); }

const char* pstm_hip_last_error(void) { return g_last_error.c_str(); }

int pstm_hip_init(int device) {
    g_last_error.clear();
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(1, "pstm_hip_init: no such device");
    HIP_TRY(hipSetDevice(device));
    if (g_state.device >= 0 && g_state.device != device) {
        return fail(4, "pstm_hip_init: library already bound to another device (call pstm_hip_release first)");
    }
    g_state.device = -1;
    return ensure_init();
}

int pstm_hip_release(void) {
    g_last_error.clear();
    if (g_state.ws) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(g_state.ws));
    }
    g_state.ws = nullptr;
    g_state.ws_size = 0;
    g_state.device = -1;
    return 0;
}

int pstm_hip_profile_reset(int enable) {
    g_last_error.clear();
    for (auto& e : g_prof.events) {
        HIP_TRY(hipEventSynchronize(e.second));
        HIP_TRY(hipEventDestroy(e.first));
        HIP_TRY(hipEventDestroy(e.second));
    }
    g_prof.events.clear();
    g_prof.on = enable != 0;
    return 0;
}

int pstm_hip_profile_read(double* total_ms, long long* launches) {
    g_last_error.clear();
    if (!total_ms || !launches) return fail(1, "output pointers must be set");
    double sum = 0.0;
    for (auto& e : g_prof.events) {
        HIP_TRY(hipEventSynchronize(e.second));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        sum += ms;
    }
    *total_ms = sum;
    *launches = (long long)g_prof.events.size();
    return 0;
}

int pstm_hip_set_tuning(int mw_cap, int batch_cap) {
    g_last_error.clear();
    if (mw_cap < 0 || mw_cap > 4 || batch_cap < 0 || batch_cap > kMaxBatch)
        return fail(1, "pstm_hip_set_tuning: mw_cap in [0, 4], batch_cap in [0, 16] (0 = default)");
    g_tune.mw_cap = mw_cap;
    g_tune.batch_cap = batch_cap;
    return 0;
}

int pstm_hip_set_pipeline(int mode) {
    g_last_error.clear();
    if (mode < -1 || mode > 1) return fail(1, "pstm_hip_set_pipeline: mode -1 (automatic), 0 (off) or 1 (on)");
    g_tune.pipe = mode;
    return 0;
}

int pstm_hip_last_launch(int* mw, int* batch, int* grid, int* items) {
    if (!mw || !batch || !grid || !items) return fail(1, "output pointers must be set");
    *mw = g_last.mw;
    *batch = g_last.batch;
    *grid = g_last.grid;
    *items = g_last.items;
    return 0;
}

int pstm_hip_stream_sync(void* stream) {
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return 0;
}

int pstm_hip_migrate_tile(const PstmMigrateArgs* args, void* stream) {
    g_last_error.clear();
    if (!args) return fail(1, "args is NULL");
    return run_migration(args, 0, args->n_traces, 1, static_cast<hipStream_t>(stream));
}

int pstm_hip_migrate_tile_noalloc(const PstmMigrateArgs* args, void* stream) {
    g_last_error.clear();
    if (!args) return fail(1, "args is NULL");
    int rc = validate(args);
    if (rc) return rc;
    if (g_prof.on)
        return fail(1, "pstm_hip_migrate_tile_noalloc: launch profiling (pstm_hip_profile_reset(1)) creates HIP "
                       "events per launch; switch it off for allocation-free calls");
    rc = ensure_init();
    if (rc) return rc;
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(4, "pstm_hip_migrate_tile_noalloc: device index out of range");
    if (!g_scratch[dev]) {
        // once per device, with that device current: the address of this
        // device's copy of the scratch array
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(3, "pstm_hip_migrate_tile_noalloc: no device address for the scratch array");
        g_scratch[dev] = static_cast<char*>(ptr);
    }
    static_assert(sizeof(scratch) == kScratchBytes, "scratch size");
    return run_migration(args, 0, args->n_traces, 1, static_cast<hipStream_t>(stream), g_scratch[dev],
                         kScratchBytes);
}

int pstm_hip_fold_tile(const PstmMigrateArgs* args, void* stream) {
    g_last_error.clear();
    int rc = validate(args);
    if (rc) return rc;
    rc = ensure_init();
    if (rc) return rc;
    if (args->n_traces == 0) return 0;
    hipLaunchKernelGGL(k_fold, dim3(args->nx * args->ny), dim3(256), 0, static_cast<hipStream_t>(stream), *args);
    HIP_TRY(hipGetLastError());
    return 0;
}

int pstm_mlp_forward_mfma_terms(const float* params, const float* x, int n, float* y, int weight_terms,
                                void* stream) {
    g_last_error.clear();
    if (!params || !x || !y || n < 0) return fail(1, "pstm_mlp_forward_mfma: bad arguments");
    if (weight_terms != 1 && weight_terms != 2) return fail(1, "pstm_mlp_forward_mfma: weight_terms must be 1 or 2");
    if (n == 0) return 0;
    int rc = ensure_init();
    if (rc) return rc;
    rc = ensure_workspace((size_t)kAVals * 4 + (size_t)kBVals * 4);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(g_state.ws);
    const int np = kAVals + kBVals;
    hipLaunchKernelGGL(k_prep_mlp, dim3((np + 255) / 256), dim3(256), 0, st, params,
                       reinterpret_cast<unsigned short*>(ws), reinterpret_cast<float*>(ws + (size_t)kAVals * 4));
    HIP_TRY(hipGetLastError());
    const long waves = ((long)n + 15) / 16;
    const auto* wa = reinterpret_cast<const bf16x8_t*>(ws);
    const auto* wb = reinterpret_cast<const f32x4_t*>(ws + (size_t)kAVals * 4);
    if (weight_terms == 2)
        hipLaunchKernelGGL(k_mlp_forward_mfma<true>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, wa, wb, x,
                           n, y);
    else
        hipLaunchKernelGGL(k_mlp_forward_mfma<false>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, wa, wb,
                           x, n, y);
    HIP_TRY(hipGetLastError());
    return 0;
}

int pstm_mlp_forward_mfma(const float* params, const float* x, int n, float* y, void* stream) {
    return pstm_mlp_forward_mfma_terms(params, x, n, y, 1, stream);
}

int pstm_hip_migrate_bins(const PstmMigrateArgs* args, const int32_t* group_start, int32_t n_groups, void* stream) {
    g_last_error.clear();
    if (!args || !group_start) return fail(1, "args and group_start must be set");
    if (n_groups < 0) return fail(1, "n_groups must be >= 0");
    if (n_groups == 0) return 0;
    for (int g = 0; g < n_groups; ++g)
        if (group_start[g + 1] < group_start[g]) return fail(1, "group_start must be non-decreasing");
    if (group_start[0] < 0 || group_start[n_groups] > args->n_traces)
        return fail(1, "group_start outside [0, n_traces]");
    // every group accumulates into the same image and fold: one gather over
    // the concatenated rows, each workgroup owning its pillars for all groups
    return run_migration(args, group_start[0], group_start[n_groups] - group_start[0], 1,
                         static_cast<hipStream_t>(stream));
}

}  // extern "C"

#else  // PSTM_GATHER_PART: one kernel group

}  // namespace

#define PSTM_GATHER_CAT2(a, b) a##b
#define PSTM_GATHER_CAT(a, b) PSTM_GATHER_CAT2(a, b)

__attribute__((visibility("hidden"))) int PSTM_GATHER_CAT(pstm_gather_part, PSTM_GATHER_PART)(
    int mw, bool count, bool lob, int interp, const void* kpv, int grid, size_t lds, hipStream_t st) {
    const KParams& kp = *static_cast<const KParams*>(kpv);
#if PSTM_GATHER_PART == 1
    (void)interp;
    return dispatch_mw<PSTM_TT_STRAIGHT, true>(mw, count, lob, kp, grid, lds, st);
#elif PSTM_GATHER_PART == 2
    (void)interp;
    return dispatch_mw<PSTM_TT_STRAIGHT, false>(mw, count, lob, kp, grid, lds, st);
#elif PSTM_GATHER_PART == 3
    (void)interp;
    return dispatch_mw<PSTM_TT_CURVED, false>(mw, count, lob, kp, grid, lds, st);
#elif PSTM_GATHER_PART == 4
    (void)interp;
    return dispatch_mw<PSTM_TT_VTI, false>(mw, count, lob, kp, grid, lds, st);
#elif PSTM_GATHER_PART == 5
    (void)interp;
    return dispatch_mw<PSTM_TT_VTI_LEARNED, false>(mw, count, lob, kp, grid, lds, st);
#elif PSTM_GATHER_PART == 7
    (void)interp;
    return dispatch_mw<PSTM_TT_CURVED_LEARNED, false>(mw, count, lob, kp, grid, lds, st);
#else
    (void)mw;
    (void)count;
    (void)lob;
    switch (interp) {
        case PSTM_INTERP_NEAREST: return launch_interp<PSTM_INTERP_NEAREST>(kp, grid, lds, st);
        case PSTM_INTERP_CUBIC: return launch_interp<PSTM_INTERP_CUBIC>(kp, grid, lds, st);
        case PSTM_INTERP_SINC4: return launch_interp<PSTM_INTERP_SINC4>(kp, grid, lds, st);
        case PSTM_INTERP_SINC8: return launch_interp<PSTM_INTERP_SINC8>(kp, grid, lds, st);
        case PSTM_INTERP_SINC16: return launch_interp<PSTM_INTERP_SINC16>(kp, grid, lds, st);
        case PSTM_INTERP_LANCZOS3: return launch_interp<PSTM_INTERP_LANCZOS3>(kp, grid, lds, st);
        case PSTM_INTERP_LANCZOS5: return launch_interp<PSTM_INTERP_LANCZOS5>(kp, grid, lds, st);
        default: return fail(1, "no gather kernel for this interpolation");
    }
#endif
}

#endif  // PSTM_GATHER_PART
