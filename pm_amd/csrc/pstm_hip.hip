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

namespace {

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
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
constexpr size_t kLdsBudget = 160 * 1024 - 2048;  // dynamic LDS; static LDS uses the rest

// per-(pillar, sample) terms, f32, built once per call in f64
//   x : traveltime slowness^2 s = 1/v^2       (curved ray: v(z) at depth)
//   y : weight constant  (spreading ? 1/v : 1) * (obliquity ? t0 : 1)
//   z : taper slope  1/(ap - ts)              (ts = ap (1 - taper))
//   w : taper offset -ts/(ap - ts)  => x_t = d*z + w, in-aperture x_t <= 1
struct KParams {
    const float* amp;            // (rows, ns)
    const int32_t* tidx;         // (ntr) or null
    const float4* geo;           // (ntr, 2): {sx, sy, rx, ry}, {mx, my, w, 0}
    const double* prow;          // AA: (ntr, rowlen) double-integrated rows
    const float4* kc;            // (npv, nt)
    const float* qk;             // (nt) tau^2 (curved: v(z)^2)
    const float2* ek;            // VTI (nev, nt): {2 eta, 1 + 2 eta}
    const float* t0k;            // (nt) t0 seconds
    const double* xax;
    const double* yax;
    float* image;                // (nx, ny, nt)
    unsigned long long* cnt_part;  // (gridDim.x) or null
    int ns, ntr, nx, ny, nt;
    int kc_pstride, ek_pstride;  // nt or 0
    int rowlen, lo;              // AA row length / index offset
    int batch;                   // traces per LDS batch
    int pair_off;                // byte offset of the pair table in LDS
    int tot_off;                 // byte offset of the f64 running totals in LDS
    float taper_off;             // x_t = d * slope + taper_off (taper_off = -(1-tf)/tf)
    int n_items, n_pgroups;      // work items = pillar groups x time windows
    int wpow;                    // 0, 1 or 2: power of 1/t in the weight
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
    const double* vrms64;        // (npv, nt)
    const double* tax64;         // (nt) ms
    const double* eta64;         // VTI: null (constant), (nt) or (npil, nt)
    int eta_pstride;             // 0 or nt (eta64 per pillar)
    int tt_req;                  // traveltime asked for (PSTM_TT_*)
    float lo_eps;                // eps_u where u >= 0 can be decided differently from f64, else -inf
    float eps_u, eps_x;          // f32 bands around the sample bounds and the aperture edge
    double tan_dip, min_ap64, max_ap64, t_start_s, dt_s, cv0, ck, eta_const;
    // learned VTI traveltime: network operand images and the t0 lattice
    const void* mlpA;            // bf16 weight operands (k_prep_mlp)
    const void* mlpB;            // f32 biases in accumulator layout
    int lat_off;                 // byte offset of the lattice in LDS
    int lat_nodes;               // lattice samples per (pair, window): 256 MW / K + 1
    int lat_shift;               // log2 K
    float lat_ifrac;             // 1 / K
    float m_tau_sc, m_v_lo, m_v_sc, m_eta_sc, m_inv_out;
};

__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }

// 0.5 (1 + cos(pi x)) for x in [0, 1]: with y = x - 1/2, cos(pi x) = -sin(pi y);
// odd Taylor series of sin(pi y) through y^11 (|pi y| <= pi/2, error < 6e-8).
__device__ __forceinline__ float taper_weight(float x) {
    const float y = x - 0.5f;
    const float z = y * y;
    float p = -7.370430945714348e-03f;        // -pi^11/11!
    p = fmaf(p, z, 8.214588661112819e-02f);   //  pi^9/9!
    p = fmaf(p, z, -5.992645293207919e-01f);  // -pi^7/7!
    p = fmaf(p, z, 2.550164039877345e+00f);   //  pi^5/5!
    p = fmaf(p, z, -5.167712780049969e+00f);  // -pi^3/3!
    p = fmaf(p, z, 3.141592653589793e+00f);   //  pi
    return fmaf(-0.5f, p * y, 0.5f);
}

// ---------------------------------------------------------------------------
// preparation kernels (f64)
// ---------------------------------------------------------------------------

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
    if (a.traveltime == PSTM_TT_CURVED) {
        const double v0 = a.curved_v0, kg = a.curved_k;
        const double tau = t_s / 2.0;
        const double z = fabs(kg) < 1e-10 ? v0 * tau : (v0 / kg) * (exp(kg * tau) - 1.0);
        const double vz = v0 + kg * z;
        wv = vz;
        if (tt_eff == PSTM_TT_STRAIGHT) {  // k ~ 0: straight ray in v0
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
    if (e < nt) {
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
        const double d = __dsqrt_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
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
__device__ __forceinline__ bool exact_decision(const KParams& P, int pil, int row, int k, double* u_out) {
    const int ix = pil / P.ny, iy = pil % P.ny;
    const double ox = P.xax[ix], oy = P.yax[iy];
    const double dxm = __dsub_rn(ox, P.mx64[row]), dym = __dsub_rn(oy, P.my64[row]);
    const double dist_mid = __dsqrt_rn(__dadd_rn(__dmul_rn(dxm, dxm), __dmul_rn(dym, dym)));
    const double v = P.vrms64[(size_t)pil * P.kc_pstride + k];
    const double t_s = P.tax64[k] / 1000.0;
    double ap = v * t_s * 0.5 * P.tan_dip;
    if (ap < P.min_ap64) ap = P.min_ap64;
    else if (ap > P.max_ap64) ap = P.max_ap64;
    const double dxs = ox - P.sx64[row], dys = oy - P.sy64[row];
    const double dxr = ox - P.rx64[row], dyr = oy - P.ry64[row];
    const double ds2 = __dadd_rn(__dmul_rn(dxs, dxs), __dmul_rn(dys, dys));
    const double dr2 = __dadd_rn(__dmul_rn(dxr, dxr), __dmul_rn(dyr, dyr));
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
        const double eta = P.eta64 ? P.eta64[(size_t)pil * P.eta_pstride + k] : P.eta_const;
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
// One wavefront evaluates the network for 32 rows with
// v_mfma_f32_32x32x8_bf16 in the transposed form D = W . X^T: A = weights
// (32 outputs x 8 inputs), B = activations (8 inputs x 32 rows), D = 32
// outputs x 32 rows.  Lane l holds row l % 32; its accumulator register r
// holds output 8 (r / 4) + 4 (l / 32) + r % 4 of the M tile, which is exactly
// the B operand layout (inputs 8 ks + 4 (l / 32) + 0..3) of K step
// ks = 4 mt + r / 4 of the next layer: activations never leave registers.
// The input and every hidden activation enter as a bf16 pair hi + lo (two
// products per K step), the weights as bf16.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf16v4_t __attribute__((ext_vector_type(4)));

constexpr int kMlpLayers = 5;
__host__ __device__ constexpr int mlp_in(int l) { return l == 0 ? PSTM_MLP_IN : PSTM_MLP_HIDDEN; }
__host__ __device__ constexpr int mlp_out(int l) { return l == kMlpLayers - 1 ? PSTM_MLP_OUT : PSTM_MLP_HIDDEN; }
__host__ __device__ constexpr int mlp_w_off(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += mlp_in(i) * mlp_out(i) + mlp_out(i);
    return off;
}
static_assert(mlp_w_off(kMlpLayers) == PSTM_MLP_PARAMS, "network layout");
// weight operands, in bf16x4 units: layer 0 [mt 2][lane 64]; layers 1-3
// [mt 2][ks 8][lane 64]; layer 4 (outputs padded to 32) [ks 8][lane 64]
constexpr int kA1 = 2 * 64, kA4 = kA1 + 3 * 16 * 64, kANum = kA4 + 8 * 64;
// biases in accumulator layout, in f32x16 units: layers 0-3 [mt 2][lane 64],
// layer 4 [lane 64]
constexpr int kB4 = 4 * 2 * 64, kBNum = kB4 + 64;

__host__ __device__ inline unsigned short bf16_bits(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    u += 0x7fffu + ((u >> 16) & 1u);  // nearest, ties to even
    return (unsigned short)(u >> 16);
}

__global__ void k_prep_mlp(const float* __restrict__ P, unsigned short* __restrict__ A, float* __restrict__ B) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kANum * 4) {
        const int v = e / 4, el = e % 4;
        int l, mt, ks, lane;
        if (v < kA1) {
            l = 0; mt = v / 64; ks = 0; lane = v % 64;
        } else if (v < kA4) {
            const int r = v - kA1;
            l = 1 + r / (16 * 64); mt = (r / (8 * 64)) % 2; ks = (r / 64) % 8; lane = r % 64;
        } else {
            l = 4; mt = 0; ks = (v - kA4) / 64; lane = (v - kA4) % 64;
        }
        const int o = mt * 32 + lane % 32, i = 8 * ks + 4 * (lane / 32) + el;
        const float w = (o < mlp_out(l) && i < mlp_in(l)) ? P[mlp_w_off(l) + o * mlp_in(l) + i] : 0.0f;
        A[e] = bf16_bits(w);
    } else if (e < kANum * 4 + kBNum * 16) {
        const int f = e - kANum * 4, v = f / 16, r = f % 16;
        int l, mt, lane;
        if (v < kB4) {
            l = v / 128; mt = (v / 64) % 2; lane = v % 64;
        } else {
            l = 4; mt = 0; lane = v - kB4;
        }
        const int o = mt * 32 + 8 * (r / 4) + 4 * (lane / 32) + r % 4;
        B[f] = o < mlp_out(l) ? P[mlp_w_off(l) + mlp_in(l) * mlp_out(l) + o] : 0.0f;
    }
}

__device__ __forceinline__ void split_bf16(float a0, float a1, float a2, float a3, bf16x4_t& hi, bf16x4_t& lo) {
    bf16v4_t h;
    h.x = (__bf16)a0; h.y = (__bf16)a1; h.z = (__bf16)a2; h.w = (__bf16)a3;
    bf16v4_t r;
    r.x = (__bf16)(a0 - (float)h.x); r.y = (__bf16)(a1 - (float)h.y);
    r.z = (__bf16)(a2 - (float)h.z); r.w = (__bf16)(a3 - (float)h.w);
    hi = __builtin_bit_cast(bf16x4_t, h);
    lo = __builtin_bit_cast(bf16x4_t, r);
}

__device__ __forceinline__ float silu(float z) { return z * frcp(1.0f + __expf(-z)); }

// SiLU of the two M tiles, re-packed as the next layer's B operands
__device__ __forceinline__ void act_pack(const f32x16_t (&acc)[2], bf16x4_t (&h)[8], bf16x4_t (&l)[8]) {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int g = 0; g < 4; ++g)
            split_bf16(silu(acc[mt][4 * g]), silu(acc[mt][4 * g + 1]), silu(acc[mt][4 * g + 2]),
                       silu(acc[mt][4 * g + 3]), h[4 * mt + g], l[4 * mt + g]);
}

__device__ __forceinline__ f32x16_t mfma2(bf16x4_t a, bf16x4_t bh, bf16x4_t bl, f32x16_t c) {
    c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bh, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bl, c, 0, 0, 0);
}

// Network outputs (y0, y1) of the row held by this lane (valid in lanes
// 0-31; lanes l and l + 32 carry the same row).  Wave-uniform call, all
// lanes active.
__device__ __forceinline__ void mlp_eval32(const bf16x4_t* __restrict__ WA, const f32x16_t* __restrict__ WB,
                                           const float (&x)[5], float& y0, float& y1) {
    const int lane = threadIdx.x & 63;
    const bool upper = lane >= 32;
    bf16x4_t h[8], l[8];
    {
        bf16x4_t xh, xl;
        split_bf16(upper ? x[4] : x[0], upper ? 0.0f : x[1], upper ? 0.0f : x[2], upper ? 0.0f : x[3], xh, xl);
        f32x16_t acc[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) acc[mt] = mfma2(WA[mt * 64 + lane], xh, xl, WB[mt * 64 + lane]);
        act_pack(acc, h, l);
    }
#pragma unroll 1
    for (int layer = 1; layer < 4; ++layer) {
        const bf16x4_t* wa = WA + kA1 + (layer - 1) * 16 * 64;
        f32x16_t acc[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            acc[mt] = WB[(layer * 2 + mt) * 64 + lane];
#pragma unroll
            for (int ks = 0; ks < 8; ++ks) acc[mt] = mfma2(wa[(mt * 8 + ks) * 64 + lane], h[ks], l[ks], acc[mt]);
        }
        act_pack(acc, h, l);
    }
    f32x16_t o = WB[kB4 + lane];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) o = mfma2(WA[kA4 + ks * 64 + lane], h[ks], l[ks], o);
    y0 = o[0];
    y1 = o[1];
}

// Batched inference through the fused path's device code: one wavefront per
// 32 rows of x (n, 5), outputs y (n, 2).
__global__ __launch_bounds__(256) void k_mlp_forward_mfma(const bf16x4_t* __restrict__ WA,
                                                          const f32x16_t* __restrict__ WB,
                                                          const float* __restrict__ x, int n, float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const long wv = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long r = wv * 32 + (lane & 31);
    float xi[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (r < n) {
#pragma unroll
        for (int i = 0; i < 5; ++i) xi[i] = x[r * 5 + i];
    }
    float y0, y1;
    mlp_eval32(WA, WB, xi, y0, y1);
    if (lane < 32 && r < n) {
        y[2 * r] = y0;
        y[2 * r + 1] = y1;
    }
}

// Traveltime of one contribution from the pair terms e/f (LDS) and the
// (pillar, sample) terms.  y1/y2 return 1/t_src, 1/t_rec for anti-aliasing.
template <int TT, bool AA>
__device__ __forceinline__ float contrib_time(const KParams& P, const float4& e, const float4& f, float q, float cs,
                                              float e2v, float e1v, float t0v, float& y1, float& y2) {
    float t;
    y1 = 0.f;
    y2 = 0.f;
    if (TT == PSTM_TT_STRAIGHT) {
        const float a1 = fmaf(e.x, cs, q);
        const float a2 = fmaf(e.y, cs, q);
        if (AA) {
            y1 = a1 > 0.f ? frsq(a1) : 0.f;
            y2 = a2 > 0.f ? frsq(a2) : 0.f;
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
    } else if (TT == PSTM_TT_VTI_LEARNED) {
        // learned legs: t^2 = tau^2 + h^2 s (1 + c), c from the lattice
        // (e2v, e1v here); 1 + c >= 0 keeps t >= t0 as the exact VTI time is
        const float a1 = fmaf(e.x * cs, fmaxf(1.0f + e2v, 0.0f), q);
        const float a2 = fmaf(e.y * cs, fmaxf(1.0f + e1v, 0.0f), q);
        t = fsqrt(a1) + fsqrt(a2);
        if (!(t0v > 0.f)) t = t0v;
    } else {
        // VTI one-way legs (anisotropy_vti.py:181-220) with s = 1/v^2:
        // t^2 = tau^2 + h^2 s - 2 eta (h^2 s)^2 / (tau^2 + (1 + 2 eta) h^2 s)
        const float hs = e.x * cs, hr = e.y * cs;
        const float ts2 = q + hs - e2v * hs * hs * frcp(fmaf(e1v, hs, q));
        const float tr2 = q + hr - e2v * hr * hr * frcp(fmaf(e1v, hr, q));
        t = fsqrt(fmaxf(ts2, 0.f)) + fsqrt(fmaxf(tr2, 0.f));
        if (!(t0v > 0.f)) t = t0v;
    }
    return t;
}

// Trace amplitude at sample position i0 + f0: linear interpolation
// (interpolation.py:68-92), or, with anti-aliasing, the triangle filter of
// half width L (from the operator time dip) on the f64 double running sum:
// a blend of the integer half widths floor(L) and floor(L)+1, each a second
// difference interpolated between i0 and i0+1.
template <bool AA, typename Row>
__device__ __forceinline__ float contrib_amp(const KParams& P, const Row* row, int i0, float f0, const float4& f,
                                             float cs, float y1, float y2) {
    if (!AA) {
        const float s0 = row[i0];
        const float s1 = row[i0 + 1];
        return fmaf(f0, s1 - s0, s0);
    } else {
        const float gx = cs * fmaf(f.x, y1, f.z * y2);
        const float gy = cs * fmaf(f.y, y1, f.w * y2);
        float L = fmaxf(fabsf(gx) * P.gdx_idt, fabsf(gy) * P.gdy_idt);
        L = fminf(fmaxf(L, 1.0f), P.lmax);
        const int lf = (int)L;
        const float lam = L - (float)lf;
        const Row* r0 = row + P.lo;
        const double am1 = r0[i0 - lf - 1], am0 = r0[i0 - lf], am1p = r0[i0 - lf + 1];
        const double c0 = r0[i0], c1 = r0[i0 + 1];
        const double b0 = r0[i0 + lf], b1 = r0[i0 + lf + 1], b2 = r0[i0 + lf + 2];
        const float dl0 = (float)((b0 - c0) - (c0 - am0));   // T_lf[i0] * lf^2
        const float dh0 = (float)((b1 - c0) - (c0 - am1));   // T_lf+1[i0] * (lf+1)^2
        const float dl1 = (float)((b1 - c1) - (c1 - am1p));  // T_lf[i0+1] * lf^2
        const float dh1 = (float)((b2 - c1) - (c1 - am0));   // T_lf+1[i0+1] * (lf+1)^2
        const float wl = (1.0f - lam) * frcp((float)(lf * lf));
        const float wh = lam * frcp((float)((lf + 1) * (lf + 1)));
        const float tr0 = fmaf(dl0, wl, dh0 * wh);
        const float tr1 = fmaf(dl1, wl, dh1 * wh);
        return fmaf(f0, tr1 - tr0, tr0);
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

template <int MW, int TT, bool AA, bool COUNT, bool LOB>
__global__ __launch_bounds__(kThreads) void k_migrate(KParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    using Row = typename std::conditional<AA, double, float>::type;
    constexpr int kPairVec = (AA || TT == PSTM_TT_CURVED) ? 2 : 1;
    Row* rows = reinterpret_cast<Row*>(smem);
    float4* pairs = reinterpret_cast<float4*>(smem + P.pair_off);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int tw = wave % kTW;
    const int pg = wave / kTW;
    const int npil = P.nx * P.ny;
    const int win = 64 * kTW * MW;
    unsigned long long cnt = 0;

    for (int item = blockIdx.x; item < P.n_items; item += gridDim.x) {
        const int pgroup = item % P.n_pgroups;
        const int kwin = item / P.n_pgroups;
        const int pil0 = pgroup * kNPil;
        const int kbase = kwin * win;

        // register-resident per-(pillar, sample) terms
        float q[MW], cs[kPW][MW], cw[kPW][MW], ci[kPW][MW];
        float e2[kPW][MW], e1[kPW][MW], t0r[MW];
        float acc[kPW][MW];
        double* tot = reinterpret_cast<double*>(smem + P.tot_off);  // [kPW*MW][kThreads], lane-private
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
                tot[(p * MW + m) * kThreads + tid] = 0.0;
                e2[p][m] = 0.0f;
                e1[p][m] = 1.0f;
                if (kv && pil < npil) {
                    const float4 c = P.kc[(size_t)pil * P.kc_pstride + k];
                    cs[p][m] = c.x;
                    cw[p][m] = c.y;
                    ci[p][m] = c.z;
                    if (TT == PSTM_TT_VTI) {
                        const float2 ee = P.ek[(size_t)pil * P.ek_pstride + k];
                        e2[p][m] = ee.x;
                        e1[p][m] = ee.y;
                    }
                } else {
                    cs[p][m] = 0.0f;
                    cw[p][m] = 0.0f;
                    ci[p][m] = INFINITY;  // x_t = +inf (d > 0) or NaN (d = 0): never in aperture
                }
            }
        }

        // sample blocks of this wave whose first t0 is <= 2 ms: only there can
        // t <= 1 ms switch the spreading/obliquity weights off (wave-uniform)
        unsigned early = 0;
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            const int kb = kbase + 64 * (tw + kTW * m);
            if (kb < P.nt && P.t0k[kb] <= 0.002f) early |= 1u << m;
        }
        if (TT == PSTM_TT_CURVED) early = (1u << MW) - 1u;  // t may be < t0 on a curved ray
        early = __builtin_amdgcn_readfirstlane(early);

        const int pt_p = tid / P.batch, pt_j = tid % P.batch;
        const bool pt_thread = tid < kNPil * P.batch;

        // learned traveltime: lattice sample below each slot and the lerp fraction
        int nidx[MW];
        float nfr = 0.0f;
        if (TT == PSTM_TT_VTI_LEARNED) {
#pragma unroll
            for (int m = 0; m < MW; ++m) nidx[m] = (64 * (tw + kTW * m) + lane) >> P.lat_shift;
            nfr = (float)(lane & ((1 << P.lat_shift) - 1)) * P.lat_ifrac;
        }

        for (int j0 = 0; j0 < P.ntr; j0 += P.batch) {
            const int nb = min(P.batch, P.ntr - j0);
            // (pillar, trace) pair table; the max-aperture test is the
            // oracle's f64 one (numba_cpu_optimized.py:167-173)
            int inap = 0;
            if (pt_thread) {
                const int pil = pil0 + pt_p;
                float4 e = make_float4(0.0f, 0.0f, INFINITY, 0.0f);
                float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (pt_j < nb && pil < npil) {
                    const double pox64 = P.xax[pil / P.ny], poy64 = P.yax[pil % P.ny];
                    const float pox = (float)(pox64 - P.xax[0]), poy = (float)(poy64 - P.yax[0]);
                    const int trow = P.tidx ? P.tidx[j0 + pt_j] : j0 + pt_j;
                    const double dx64 = __dsub_rn(pox64, P.mx64[trow]);
                    const double dy64 = __dsub_rn(poy64, P.my64[trow]);
                    const double dm64 = __dsqrt_rn(__dadd_rn(__dmul_rn(dx64, dx64), __dmul_rn(dy64, dy64)));
                    if (!(dm64 > P.max_ap64)) {
                        inap = 1;
                        const float4 g0 = P.geo[2 * (j0 + pt_j)];
                        const float4 g1 = P.geo[2 * (j0 + pt_j) + 1];
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
                        } else {
                            e = make_float4(ds2, dr2, dm, g1.z);
                            f = make_float4(dxs, dys, dxr, dyr);
                        }
                    }
                }
                pairs[(pt_p * P.batch + pt_j) * kPairVec] = e;
                if (kPairVec == 2) pairs[(pt_p * P.batch + pt_j) * kPairVec + 1] = f;
            }
            if (!__syncthreads_or(inap)) continue;

            if (TT == PSTM_TT_VTI_LEARNED) {
                // network on the t0 lattice of every in-aperture pair of the
                // batch, 32 (pair, lattice sample) rows per wavefront step
                const int nn = P.lat_nodes;
                const int R = kNPil * P.batch * nn;
                float2* lat = reinterpret_cast<float2*>(smem + P.lat_off);
                const float dt_last = P.nt > 1 ? P.t0k[P.nt - 1] - P.t0k[P.nt - 2] : 0.0f;
                for (int c0 = wave * 32; c0 < R; c0 += kWaves * 32) {
                    const int r = c0 + (lane & 31);
                    bool valid = false;
                    float x[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                    if (r < R) {
                        const int pp = r / nn, node = r - pp * nn;
                        const float4 e = pairs[pp * kPairVec];
                        if (e.z != INFINITY) {
                            valid = true;
                            const int pil = pil0 + pp / P.batch;
                            const int k = kbase + (node << P.lat_shift);
                            const int kk = min(k, P.nt - 1);
                            const float t0 = P.t0k[kk] + (float)(k - kk) * dt_last;
                            const float tau = 0.5f * t0;
                            const float q = tau * tau;
                            const float sl = P.kc[(size_t)pil * P.kc_pstride + kk].x;  // 1/v^2
                            const float eta = 0.5f * P.ek[(size_t)pil * P.ek_pstride + kk].x;
                            const float hs = e.x * sl, hr = e.y * sl;
                            const float ts2 = q + hs, tr2 = q + hr;
                            const float sn_s = ts2 > 0.0f ? fsqrt(hs / ts2) : 0.0f;
                            const float sn_r = tr2 > 0.0f ? fsqrt(hr / tr2) : 0.0f;
                            const float v = frsq(sl);
                            x[0] = fmaf(tau, P.m_tau_sc, -1.0f);
                            x[1] = fmaf(2.0f, sn_s, -1.0f);
                            x[2] = fmaf(2.0f, sn_r, -1.0f);
                            x[3] = fmaf(v - P.m_v_lo, P.m_v_sc, -1.0f);
                            x[4] = fmaf(eta, P.m_eta_sc, -1.0f);
                        }
                    }
                    if (!__any(valid)) continue;  // wave-uniform
                    float y0, y1;
                    mlp_eval32(static_cast<const bf16x4_t*>(P.mlpA), static_cast<const f32x16_t*>(P.mlpB), x, y0, y1);
                    if (lane < 32 && valid) lat[r] = make_float2(y0 * P.m_inv_out, y1 * P.m_inv_out);
                }
            }

            // stage the batch of trace rows in LDS (coalesced)
            if (AA) {
                const double* src = P.prow + (size_t)j0 * P.rowlen;
                double* dst = reinterpret_cast<double*>(rows);
                const int n = nb * P.rowlen;
                for (int e = tid; e < n; e += kThreads) dst[e] = src[e];
            } else {
                float* dst = reinterpret_cast<float*>(rows);
                for (int r = 0; r < nb; ++r) {
                    const int row = P.tidx ? P.tidx[j0 + r] : j0 + r;
                    const float* src = P.amp + (size_t)row * P.ns;
                    for (int e = tid; e < P.ns; e += kThreads) dst[r * P.rowlen + e] = src[e];
                }
            }
            __syncthreads();

            for (int jj = 0; jj < nb; ++jj) {
                const Row* row = rows + (size_t)jj * P.rowlen;
                bool any_near = false;  // some slot within eps of a threshold (rare)
#pragma unroll
                for (int p = 0; p < kPW; ++p) {
                    const int pi = (pg * kPW + p) * P.batch + jj;
                    const float4 e = pairs[pi * kPairVec];
                    if (e.z == INFINITY) continue;  // pair outside max aperture (wave-uniform)
                    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (kPairVec == 2) f = pairs[pi * kPairVec + 1];
                    // Every slot of the pair is evaluated without per-slot branches so
                    // that the MW independent chains (sqrt, LDS reads) interleave;
                    // slots outside the aperture or the trace read a safe sample and
                    // add zero.
                    float xt[MW];
                    bool any_in = false;
#pragma unroll
                    for (int m = 0; m < MW; ++m) {
                        xt[m] = fmaf(e.z, ci[p][m], P.taper_off);
                        any_in |= xt[m] <= 1.0f + P.eps_x;
                    }
                    if (!any_in) continue;  // time-dependent aperture excludes every slot
                    float t[MW], u[MW], y1[MW], y2[MW];
                    bool ok[MW];
                    if (TT == PSTM_TT_VTI_LEARNED) {
                        const float2* lp = reinterpret_cast<const float2*>(smem + P.lat_off) + (size_t)pi * P.lat_nodes;
#pragma unroll
                        for (int m = 0; m < MW; ++m) {
                            const float2 c0 = lp[nidx[m]], c1 = lp[nidx[m] + 1];
                            e2[p][m] = fmaf(nfr, c1.x - c0.x, c0.x);
                            e1[p][m] = fmaf(nfr, c1.y - c0.y, c0.y);
                        }
                    }
#pragma unroll
                    for (int m = 0; m < MW; ++m) {
                        t[m] = contrib_time<TT, AA>(P, e, f, q[m], cs[p][m], e2[p][m], e1[p][m], t0r[m], y1[m],
                                                    y2[m]);
                        u[m] = fmaf(t[m], P.idt, P.uoff);
                        // u < umax - eps: inside; [umax - eps, umax + eps): decided in f64
                        const bool in_ap = xt[m] <= ((COUNT && TT != PSTM_TT_VTI_LEARNED) ? 1.0f - P.eps_x : 1.0f);
                        const bool below = u[m] < P.umax_lo;
                        bool near = in_ap && !below && u[m] < P.umax_hi;
                        if (LOB) near = near || (in_ap && u[m] < P.lo_eps && u[m] > -P.lo_eps);
                        if (COUNT) near = near || (fabsf(xt[m] - 1.0f) < P.eps_x);
                        // the learned time has no f64 counterpart to decide against
                        if (TT == PSTM_TT_VTI_LEARNED) near = false;
                        any_near |= near;
                        ok[m] = in_ap && (TT == PSTM_TT_VTI_LEARNED ? u[m] < P.umax : below) && !near;
                        if (LOB) ok[m] = ok[m] && u[m] >= 0.0f;
                    }
                    float tw[MW];
                    bool any_taper = false;
#pragma unroll
                    for (int m = 0; m < MW; ++m) {
                        tw[m] = e.w;
                        any_taper |= ok[m] && xt[m] > 0.0f;
                    }
                    if (any_taper) {
#pragma unroll
                        for (int m = 0; m < MW; ++m)
                            if (xt[m] > 0.0f) tw[m] = e.w * taper_weight(fminf(xt[m], 1.0f));
                    }
#pragma unroll
                    for (int m = 0; m < MW; ++m) {
                        const float us = ok[m] ? u[m] : 0.0f;
                        const int i0 = (int)us;
                        const float f0 = __builtin_amdgcn_fractf(us);
                        const float amp = contrib_amp<AA>(P, row, i0, f0, f, cs[p][m], y1[m], y2[m]);
                        float w = ok[m] ? tw[m] : 0.0f;
                        if (P.wpow != 0) {
                            const float r = frcp(t[m]);
                            const float wf = cw[p][m] * (P.wpow == 2 ? r * r : r);
                            // t > 1 ms is certain unless the block holds t0 <= 2 ms (early)
                            if (early & (1u << m)) w = t[m] > 0.001f ? w * wf : w;
                            else w *= wf;
                        }
                        acc[p][m] = fmaf(amp, w, acc[p][m]);
                        if (COUNT) cnt += ok[m] ? 1 : 0;
                    }
                }
                // slots near a threshold: recompute which, decide in f64 as the
                // oracle does (rare; the slots were left out above)
                if (TT != PSTM_TT_VTI_LEARNED && any_near) {
#pragma unroll
                    for (int p = 0; p < kPW; ++p) {
                        const int pi = (pg * kPW + p) * P.batch + jj;
                        const float4 e = pairs[pi * kPairVec];
                        if (e.z == INFINITY) continue;
                        float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (kPairVec == 2) f = pairs[pi * kPairVec + 1];
#pragma unroll
                        for (int m = 0; m < MW; ++m) {
                            const float xt = fmaf(e.z, ci[p][m], P.taper_off);
                            float y1, y2;
                            const float t = contrib_time<TT, AA>(P, e, f, q[m], cs[p][m], e2[p][m], e1[p][m],
                                                                 t0r[m], y1, y2);
                            const float u = fmaf(t, P.idt, P.uoff);
                            const bool in_ap = xt <= (COUNT ? 1.0f - P.eps_x : 1.0f);
                            bool near = in_ap && u >= P.umax_lo && u < P.umax_hi;
                            if (LOB) near = near || (in_ap && u < P.lo_eps && u > -P.lo_eps);
                            if (COUNT) near = near || (fabsf(xt - 1.0f) < P.eps_x);
                            if (!near) continue;
                            const int pil = pil0 + pg * kPW + p;
                            const int k = kbase + 64 * (tw + kTW * m) + lane;
                            const int trow = P.tidx ? P.tidx[j0 + jj] : j0 + jj;
                            double u64;
                            if (exact_decision<TT>(P, pil, trow, k, &u64)) {
                                const int i0 = (int)u64;
                                const float f0 = (float)(u64 - (double)i0);
                                acc[p][m] = fmaf(contrib_amp<AA>(P, row, i0, f0, f, cs[p][m], y1, y2),
                                                 contrib_weight(P, e.w, xt, t, cw[p][m]), acc[p][m]);
                                if (COUNT) ++cnt;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < kPW; ++p)
#pragma unroll
                for (int m = 0; m < MW; ++m) {
                    tot[(p * MW + m) * kThreads + tid] += (double)acc[p][m];
                    acc[p][m] = 0.0f;
                }
            __syncthreads();
        }

        // the workgroup is the only writer of these cells
#pragma unroll
        for (int m = 0; m < MW; ++m) {
            const int k = kbase + 64 * (tw + kTW * m) + lane;
            if (k >= P.nt) continue;
#pragma unroll
            for (int p = 0; p < kPW; ++p) {
                const int pil = pil0 + pg * kPW + p;
                if (pil < npil) {
                    float* dst = P.image + (size_t)pil * P.nt + k;
                    *dst = *dst + (float)tot[(p * MW + m) * kThreads + tid];
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
    if (a->traveltime < PSTM_TT_STRAIGHT || a->traveltime > PSTM_TT_VTI_LEARNED) return fail(1, "invalid traveltime");
    if (a->aa_mode != PSTM_AA_NONE && a->aa_mode != PSTM_AA_TRIANGLE) return fail(1, "invalid aa_mode");
    if (a->aa_mode == PSTM_AA_TRIANGLE && a->traveltime != PSTM_TT_STRAIGHT)
        return fail(1, "triangle anti-aliasing is implemented for the straight-ray traveltime only");
    if (a->aa_mode == PSTM_AA_TRIANGLE && (a->aa_max_half_width < 1 || a->aa_max_half_width > 64))
        return fail(1, "aa_max_half_width must be in [1, 64]");
    if (a->traveltime == PSTM_TT_VTI || a->traveltime == PSTM_TT_VTI_LEARNED) {
        if (a->eta_mode < PSTM_ETA_CONST || a->eta_mode > PSTM_ETA_3D) return fail(1, "invalid eta_mode");
        if (a->eta_mode != PSTM_ETA_CONST && !a->eta) return fail(1, "eta pointer must be set");
    }
    if (a->traveltime == PSTM_TT_VTI_LEARNED) {
        const int K = a->mlp_node_spacing;
        if (!a->mlp_params) return fail(1, "mlp_params must be set for the learned traveltime");
        if (K != 8 && K != 16 && K != 32 && K != 64) return fail(1, "mlp_node_spacing must be 8, 16, 32 or 64");
        if (!(a->mlp_tau_max > 0.0) || !(a->mlp_v_hi > a->mlp_v_lo) || !(a->mlp_eta_max > 0.0) ||
            !(a->mlp_out_scale > 0.0))
            return fail(1, "learned traveltime encoding: need tau_max > 0, v_hi > v_lo, eta_max > 0, out_scale > 0");
    }
    if (a->traveltime == PSTM_TT_CURVED && !(a->curved_v0 > 0.0))
        return fail(1, "curved_v0 must be positive");
    if (!(a->dt_in_ms > 0.0)) return fail(1, "dt_in_ms must be positive");
    if (!(a->max_aperture_m > 0.0)) return fail(1, "max_aperture_m must be positive");
    if ((long long)a->nx * a->ny * a->nt > (1LL << 40)) return fail(1, "tile too large");
    return 0;
}

template <int MW, int TT, bool AA, bool COUNT, bool LOB>
int launch_migrate(const KParams& kp, int grid, size_t lds, hipStream_t st) {
    auto fn = k_migrate<MW, TT, AA, COUNT, LOB>;
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (g_prof.on) {
        HIP_TRY(hipEventCreate(&ev0));
        HIP_TRY(hipEventCreate(&ev1));
        HIP_TRY(hipEventRecord(ev0, st));
    }
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), lds, st, kp);
    HIP_TRY(hipGetLastError());
    if (g_prof.on) {
        HIP_TRY(hipEventRecord(ev1, st));
        g_prof.events.emplace_back(ev0, ev1);
    }
    return 0;
}

template <int TT, bool AA, bool COUNT, bool LOB>
int dispatch_mw3(int mw, const KParams& kp, int grid, size_t lds, hipStream_t st) {
    switch (mw) {
        case 1: return launch_migrate<1, TT, AA, COUNT, LOB>(kp, grid, lds, st);
        case 2: return launch_migrate<2, TT, AA, COUNT, LOB>(kp, grid, lds, st);
        case 3: return launch_migrate<3, TT, AA, COUNT, LOB>(kp, grid, lds, st);
        default: return launch_migrate<4, TT, AA, COUNT, LOB>(kp, grid, lds, st);
    }
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

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// shared by migrate_tile and migrate_bins: `ntr` rows starting at row
// offset `row0` of args->trace_index (or of the block).
int run_migration(const PstmMigrateArgs* in, int row0, int ntr, int do_fold, hipStream_t st) {
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
    const int nt = a.nt, ns = a.n_samples;
    const int npil = a.nx * a.ny;
    const int npv = a.vel_mode == PSTM_VEL_3D ? npil : 1;
    const int nev = a.eta_mode == PSTM_ETA_3D ? npil : 1;
    int tt = a.traveltime;
    if (tt == PSTM_TT_CURVED && fabs(a.curved_k) < 1e-10) tt = PSTM_TT_STRAIGHT;
    const bool aa = a.aa_mode == PSTM_AA_TRIANGLE;
    const int lmax = aa ? a.aa_max_half_width : 0;
    const int lo = aa ? lmax + 1 : 0;
    const int rowlen = aa ? ns + 2 * lmax + 2 : ns;
    const size_t elem = aa ? sizeof(double) : sizeof(float);
    const int pair_vec = (aa || tt == PSTM_TT_CURVED) ? 2 : 1;
    const bool learned = tt == PSTM_TT_VTI_LEARNED;
    const bool vti_like = tt == PSTM_TT_VTI || learned;

    // batch size from the LDS budget
    // time windows
    const int nwin = (nt + 1023) / 1024;
    const int per_win = (nt + nwin - 1) / nwin;
    int mw_cap = g_tune.mw_cap > 0 ? g_tune.mw_cap : env_int("PSTM_HIP_MW");
    if (mw_cap <= 0) mw_cap = 4;
    mw_cap = std::min(4, mw_cap);
    int mw = std::min(mw_cap, std::max(1, (per_win + 255) / 256));
    // batch size from the LDS budget: trace rows + pair table + f64 totals
    const size_t tot_bytes = (size_t)kPW * mw * kThreads * sizeof(double);
    const int lat_nodes = learned ? 256 * mw / a.mlp_node_spacing + 1 : 0;
    auto lds_for = [&](int b) {
        return align_up(align_up((size_t)b * rowlen * elem, 16) + (size_t)kNPil * b * pair_vec * 16, 16) + tot_bytes +
               align_up((size_t)kNPil * b * lat_nodes * 8, 16);
    };
    int batch_cap = g_tune.batch_cap > 0 ? g_tune.batch_cap : env_int("PSTM_HIP_BATCH");
    int batch = batch_cap > 0 ? std::min(batch_cap, kMaxBatch) : kMaxBatch;
    while (batch > 1 && lds_for(batch) > kLdsBudget) --batch;
    const size_t lds = lds_for(batch);
    if (lds > kLdsBudget) return fail(5, "trace too long for one LDS row (n_samples too large)");
    const int win = 256 * mw;
    const int nwin2 = (nt + win - 1) / win;
    const int n_pgroups = (npil + kNPil - 1) / kNPil;
    const long n_items = (long)n_pgroups * nwin2;
    if (n_items > (1L << 30)) return fail(1, "tile too large");
    const int occupancy = lds <= kLdsBudget / 2 ? 2 : 1;  // workgroups per CU the LDS allows
    const int grid = (int)std::min<long>(n_items, (long)g_state.n_cu * occupancy);

    g_last.mw = mw;
    g_last.batch = batch;
    g_last.grid = grid;
    g_last.items = (int)n_items;

    // workspace layout
    size_t off = 0;
    const size_t o_geo = off; off = align_up(off + (size_t)ntr * 32, 256);
    const size_t o_kc = off; off = align_up(off + (size_t)npv * nt * 16, 256);
    const size_t o_qk = off; off = align_up(off + (size_t)nt * 4, 256);
    const size_t o_t0 = off; off = align_up(off + (size_t)nt * 4, 256);
    const size_t o_ek = off; off = align_up(off + (vti_like ? (size_t)nev * nt * 8 : 0), 256);
    const size_t o_mlpA = off; off = align_up(off + (learned ? (size_t)kANum * 8 : 0), 256);
    const size_t o_mlpB = off; off = align_up(off + (learned ? (size_t)kBNum * 64 : 0), 256);
    const size_t o_cnt = off; off = align_up(off + (size_t)grid * 8, 256);
    const size_t o_prow = off; off = align_up(off + (aa ? (size_t)ntr * rowlen * 8 : 0), 256);
    rc = ensure_workspace(off);
    if (rc) return rc;
    char* ws = static_cast<char*>(g_state.ws);
    float4* geo = reinterpret_cast<float4*>(ws + o_geo);
    float4* kc = reinterpret_cast<float4*>(ws + o_kc);
    float* qk = reinterpret_cast<float*>(ws + o_qk);
    float* t0k = reinterpret_cast<float*>(ws + o_t0);
    float2* ek = reinterpret_cast<float2*>(ws + o_ek);
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(ws + o_cnt);
    double* prow = reinterpret_cast<double*>(ws + o_prow);

    hipLaunchKernelGGL(k_prep_geometry, dim3((ntr + 255) / 256), dim3(256), 0, st, a, geo);
    HIP_TRY(hipGetLastError());
    const double tan_dip = tan(a.max_dip_deg * M_PI / 180.0);
    const long nkc = (long)npv * nt;
    hipLaunchKernelGGL(k_prep_time_terms, dim3((unsigned)((nkc + 255) / 256)), dim3(256), 0, st, a, tt, npv, tan_dip,
                       kc, qk, t0k);
    HIP_TRY(hipGetLastError());
    if (learned) {
        const int n = kANum * 4 + kBNum * 16;
        hipLaunchKernelGGL(k_prep_mlp, dim3((n + 255) / 256), dim3(256), 0, st, a.mlp_params,
                           reinterpret_cast<unsigned short*>(ws + o_mlpA), reinterpret_cast<float*>(ws + o_mlpB));
        HIP_TRY(hipGetLastError());
    }
    if (vti_like) {
        const long nek = (long)nev * nt;
        hipLaunchKernelGGL(k_prep_eta, dim3((unsigned)((nek + 255) / 256)), dim3(256), 0, st, a, nev, ek);
        HIP_TRY(hipGetLastError());
    }
    if (aa) {
        hipLaunchKernelGGL(k_integrate, dim3(ntr), dim3(256), 0, st, a, rowlen, lo, prow);
        HIP_TRY(hipGetLastError());
    }
    if (do_fold) {
        hipLaunchKernelGGL(k_fold, dim3(npil), dim3(256), 0, st, a);
        HIP_TRY(hipGetLastError());
    }

    KParams kp;
    kp.amp = a.amplitudes;
    kp.tidx = a.trace_index;
    kp.geo = geo;
    kp.prow = prow;
    kp.kc = kc;
    kp.qk = qk;
    kp.ek = ek;
    kp.t0k = t0k;
    kp.xax = a.x_axis;
    kp.yax = a.y_axis;
    kp.image = a.image;
    kp.cnt_part = a.n_effective ? cnt : nullptr;
    kp.ns = ns;
    kp.ntr = ntr;
    kp.nx = a.nx;
    kp.ny = a.ny;
    kp.nt = nt;
    kp.kc_pstride = npv > 1 ? nt : 0;
    kp.ek_pstride = nev > 1 ? nt : 0;
    kp.rowlen = rowlen;
    kp.lo = lo;
    kp.batch = batch;
    kp.pair_off = (int)align_up((size_t)batch * rowlen * elem, 16);
    kp.tot_off = (int)align_up((size_t)kp.pair_off + (size_t)kNPil * batch * pair_vec * 16, 16);
    kp.taper_off = a.taper_fraction > 0.0 ? (float)(-(1.0 - a.taper_fraction) / a.taper_fraction) : -1e30f;
    kp.n_items = (int)n_items;
    kp.n_pgroups = n_pgroups;
    kp.wpow = (a.apply_spreading ? 1 : 0) + (a.apply_obliquity ? 1 : 0);
    const double dt_s = a.dt_in_ms / 1000.0;
    kp.idt = (float)(1.0 / dt_s);
    kp.uoff = (float)(-(a.t_start_in_ms / 1000.0) / dt_s);
    kp.umax = (float)(ns - 1);
    kp.max_ap = (float)a.max_aperture_m;
    kp.gdx_idt = (float)(a.grid_dx / dt_s);
    kp.gdy_idt = (float)(a.grid_dy / dt_s);
    kp.lmax = (float)lmax;
    kp.ck2 = (float)(a.curved_k * a.curved_k);
    kp.ik_ln2 = (float)(tt == PSTM_TT_CURVED ? M_LN2 / a.curved_k : 0.0);
    kp.v0 = (float)a.curved_v0;
    kp.t_floor = a.curved_k > 0.0 ? 0.0f : -INFINITY;
    kp.sx64 = a.source_x;
    kp.sy64 = a.source_y;
    kp.rx64 = a.receiver_x;
    kp.ry64 = a.receiver_y;
    kp.mx64 = a.midpoint_x;
    kp.my64 = a.midpoint_y;
    kp.vrms64 = a.vrms;
    kp.tax64 = a.t_axis_ms;
    kp.eta64 = (vti_like && a.eta_mode != PSTM_ETA_CONST) ? a.eta : nullptr;
    kp.eta_pstride = a.eta_mode == PSTM_ETA_3D ? nt : 0;
    kp.tt_req = a.traveltime;
    // u >= 0 can only fail or be decided differently from f64 when t can be
    // at or below t_start: t_start > 0, a curved ray with k < 0, or VTI
    // VTI returns t = t0 for t0 <= 0 (negative output times give u < 0)
    const bool lo_check = a.t_start_in_ms > 0.0 || (tt == PSTM_TT_CURVED && a.curved_k < 0.0) || vti_like;
    kp.eps_u = (float)(2e-3 + 1e-5 * (double)(ns - 1));
    kp.lo_eps = lo_check ? kp.eps_u : -INFINITY;
    kp.umax_lo = kp.umax - kp.eps_u;
    kp.umax_hi = kp.umax + kp.eps_u;
    kp.eps_x = a.n_effective ? 1e-4f : 0.0f;
    kp.tan_dip = tan_dip;
    kp.min_ap64 = a.min_aperture_m;
    kp.max_ap64 = a.max_aperture_m;
    kp.t_start_s = a.t_start_in_ms / 1000.0;
    kp.dt_s = a.dt_in_ms / 1000.0;
    kp.cv0 = a.curved_v0;
    kp.ck = a.curved_k;
    kp.eta_const = a.eta_constant;
    kp.mlpA = learned ? ws + o_mlpA : nullptr;
    kp.mlpB = learned ? ws + o_mlpB : nullptr;
    kp.lat_off = (int)(kp.tot_off + tot_bytes);
    kp.lat_nodes = lat_nodes;
    kp.lat_shift = learned ? __builtin_ctz((unsigned)a.mlp_node_spacing) : 0;
    kp.lat_ifrac = learned ? 1.0f / (float)a.mlp_node_spacing : 0.0f;
    kp.m_tau_sc = learned ? (float)(2.0 / a.mlp_tau_max) : 0.0f;
    kp.m_v_lo = (float)a.mlp_v_lo;
    kp.m_v_sc = learned ? (float)(2.0 / (a.mlp_v_hi - a.mlp_v_lo)) : 0.0f;
    kp.m_eta_sc = learned ? (float)(2.0 / a.mlp_eta_max) : 0.0f;
    kp.m_inv_out = learned ? (float)(1.0 / a.mlp_out_scale) : 0.0f;

    const bool count = a.n_effective != nullptr;
    if (tt == PSTM_TT_STRAIGHT && aa) rc = dispatch_mw<PSTM_TT_STRAIGHT, true>(mw, count, lo_check, kp, grid, lds, st);
    else if (tt == PSTM_TT_STRAIGHT) rc = dispatch_mw<PSTM_TT_STRAIGHT, false>(mw, count, lo_check, kp, grid, lds, st);
    else if (tt == PSTM_TT_CURVED) rc = dispatch_mw<PSTM_TT_CURVED, false>(mw, count, lo_check, kp, grid, lds, st);
    else if (tt == PSTM_TT_VTI) rc = dispatch_mw<PSTM_TT_VTI, false>(mw, count, lo_check, kp, grid, lds, st);
    else rc = dispatch_mw<PSTM_TT_VTI_LEARNED, false>(mw, count, lo_check, kp, grid, lds, st);
    if (rc) return rc;
    if (a.n_effective) {
        hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(256), 0, st, (const unsigned long long*)cnt, grid,
                           a.n_effective);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

}  // namespace

extern "C" {

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

int pstm_mlp_forward_mfma(const float* params, const float* x, int n, float* y, void* stream) {
    g_last_error.clear();
    if (!params || !x || !y || n < 0) return fail(1, "pstm_mlp_forward_mfma: bad arguments");
    if (n == 0) return 0;
    int rc = ensure_init();
    if (rc) return rc;
    rc = ensure_workspace((size_t)kANum * 8 + (size_t)kBNum * 64);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(g_state.ws);
    const int np = kANum * 4 + kBNum * 16;
    hipLaunchKernelGGL(k_prep_mlp, dim3((np + 255) / 256), dim3(256), 0, st, params,
                       reinterpret_cast<unsigned short*>(ws), reinterpret_cast<float*>(ws + (size_t)kANum * 8));
    HIP_TRY(hipGetLastError());
    const long waves = ((long)n + 31) / 32;
    hipLaunchKernelGGL(k_mlp_forward_mfma, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st,
                       reinterpret_cast<const bf16x4_t*>(ws),
                       reinterpret_cast<const f32x16_t*>(ws + (size_t)kANum * 8), x, n, y);
    HIP_TRY(hipGetLastError());
    return 0;
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
