// pstm_mlp.hip — traveltime surrogate: a 5-64-64-64-64-2 SiLU MLP trained and
// evaluated by hand-written gfx950 kernels.
//
// The network maps the encoded (tau, s_s, s_r, v, eta) of a (pillar, trace,
// t0) triple (s = sine of a leg's straight-ray angle) to out_scale times the
// anisotropic factor c of each one-way leg, t_leg^2 = tau^2 + (d/v)^2 (1 + c).
// The exact targets come from the one-way VTI times of
// pstm/algorithm/anisotropy_vti.py:181-220 evaluated in f64 (k_vti_targets).
//
// Parameters live in one f32 buffer (layout below).  Training is f32:
// per-layer forward kernels keep pre-activations for the backward pass, the
// loss is the mean squared error over both outputs, and Adam updates every
// parameter in place.  The migration kernel evaluates the same network in
// BF16 with f32 accumulation (pstm_hip.hip).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pstm_hip.h"

namespace {

constexpr int kIn = PSTM_MLP_IN;       // 5
constexpr int kH = PSTM_MLP_HIDDEN;    // 64
constexpr int kOut = PSTM_MLP_OUT;     // 2
constexpr int kLayers = 5;             // 4 hidden + output

// parameter layout: for each layer l, W_l (out x in, row-major) then b_l (out)
__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kIn : kH; }
__host__ __device__ constexpr int layer_out(int l) { return l == kLayers - 1 ? kOut : kH; }
__host__ __device__ constexpr int layer_w_off(int l) {
    int off = 0;
    for (int i = 0; i < l; ++i) off += layer_in(i) * layer_out(i) + layer_out(i);
    return off;
}
__host__ __device__ constexpr int layer_b_off(int l) { return layer_w_off(l) + layer_in(l) * layer_out(l); }
constexpr int kParams = layer_w_off(kLayers);
static_assert(kParams == PSTM_MLP_PARAMS, "parameter count");

thread_local std::string g_err;
int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}
#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) return fail(3, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

// f32 -> nearest bf16 (ties to even), returned as f32
__device__ __forceinline__ float bf16r(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// a carried as a pair of bf16 values, hi + lo (about 16 significant bits)
__device__ __forceinline__ float bf16pair(float a) {
    const float hi = bf16r(a);
    return hi + bf16r(a - hi);
}

// Precision of the forward pass.  Q = false: f32.  Q = true: the numerics of
// the migration kernel's fused inference, i.e. weights rounded to bf16, the
// network input and every hidden activation carried as a bf16 pair (hi + lo,
// two MFMA products per layer), f32 accumulation and biases (the inference
// carries its values times -log2(e), so its pairs round a scaled value: the
// same relative precision, other last bits).  Training with Q = true is
// quantisation-aware (straight-through gradients).
bool g_q = false;

// Z[s, o] = b[o] + sum_i X[s, i] W[o, i];  A = silu(Z) (or Z for the output layer)
template <bool ACT, bool Q>
__global__ void k_linear_fwd(const float* __restrict__ P, int l, const float* __restrict__ X, int n,
                             float* __restrict__ Z, float* __restrict__ A) {
    const int in = layer_in(l), out = layer_out(l);
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * out) return;
    const int s = (int)(e / out), o = (int)(e % out);
    const float* W = P + layer_w_off(l) + o * in;
    const float* x = X + (size_t)s * in;
    float z = P[layer_b_off(l) + o];
    for (int i = 0; i < in; ++i) {
        const float xv = (Q && l == 0) ? bf16pair(x[i]) : x[i];
        z = fmaf(xv, Q ? bf16r(W[i]) : W[i], z);
    }
    if (Z) Z[e] = z;
    const float a = ACT ? z * sigmoidf_(z) : z;
    A[e] = (Q && ACT) ? bf16pair(a) : a;
}

template <bool ACT>
void launch_fwd(unsigned nb, hipStream_t st, const float* P, int l, const float* X, int n, float* Z, float* A) {
    if (g_q)
        hipLaunchKernelGGL((k_linear_fwd<ACT, true>), dim3(nb), dim3(256), 0, st, P, l, X, n, Z, A);
    else
        hipLaunchKernelGGL((k_linear_fwd<ACT, false>), dim3(nb), dim3(256), 0, st, P, l, X, n, Z, A);
}

// dZ = dA * silu'(Z) (in place in dA), output layer: dZ = dA
__global__ void k_silu_bwd(const float* __restrict__ Z, float* __restrict__ dA, long n) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float z = Z[e];
    const float sg = sigmoidf_(z);
    dA[e] *= sg * (1.0f + z * (1.0f - sg));
}

// dX[s, i] = sum_o dZ[s, o] W[o, i]  (W rounded to bf16 when Q)
template <bool Q>
__global__ void k_linear_bwd_x(const float* __restrict__ P, int l, const float* __restrict__ dZ, int n,
                               float* __restrict__ dX) {
    const int in = layer_in(l), out = layer_out(l);
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * in) return;
    const int s = (int)(e / in), i = (int)(e % in);
    const float* W = P + layer_w_off(l);
    const float* dz = dZ + (size_t)s * out;
    float acc = 0.0f;
    for (int o = 0; o < out; ++o) acc = fmaf(dz[o], Q ? bf16r(W[o * in + i]) : W[o * in + i], acc);
    dX[e] = acc;
}

// dW[o, i] = sum_s dZ[s, o] X[s, i], db[o] = sum_s dZ[s, o]: one block per
// (o, i) or (o, bias), block reduction over the samples (no atomics)
__global__ __launch_bounds__(256) void k_linear_bwd_w(int l, const float* __restrict__ dZ,
                                                      const float* __restrict__ X, int n,
                                                      float* __restrict__ G) {
    const int in = layer_in(l), out = layer_out(l);
    const int e = blockIdx.x;  // 0 .. out*(in+1)-1
    const int o = e / (in + 1), i = e % (in + 1);
    float acc = 0.0f;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        const float xv = i < in ? X[(size_t)s * in + i] : 1.0f;
        acc = fmaf(dZ[(size_t)s * out + o], xv, acc);
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (i < in) G[layer_w_off(l) + o * in + i] = red[0];
        else G[layer_b_off(l) + o] = red[0];
    }
}

// MSE over n x kOut outputs: dY = 2 (Y - T) / (n kOut); per-block partial losses
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int n,
                                             float* __restrict__ dY, float* __restrict__ part) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)n * kOut;
    float l = 0.0f;
    if (e < tot) {
        const float d = Y[e] - T[e];
        dY[e] = 2.0f * d / (float)tot;
        l = d * d;
    }
    __shared__ float red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_loss(const float* __restrict__ part, int nb, float scale, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

// Adam (Kingma & Ba 2015) with bias correction, in place
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, int np, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const float g = G[e];
    const float m = b1 * M[e] + (1.0f - b1) * g;
    const float v = b2 * V[e] + (1.0f - b2) * g * g;
    M[e] = m;
    V[e] = v;
    P[e] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

// exact one-way VTI times (anisotropy_vti.py:181-220), f64
__global__ void k_vti_targets(const double* __restrict__ tau, const double* __restrict__ h,
                              const double* __restrict__ v, const double* __restrict__ eta, long n,
                              double* __restrict__ t_out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double t0 = tau[e], hh = h[e], vn = v[e], et = eta[e];
    double t;
    if (vn <= 0 || t0 <= 0) {
        t = t0;
    } else {
        const double h2 = hh * hh, v2 = vn * vn, t02 = t0 * t0;
        const double t2h = t02 + h2 / v2;
        const double den = v2 * (t02 * v2 + (1.0 + 2.0 * et) * h2);
        t = fabs(den) < 1e-20 ? sqrt(fmax(t2h, 0.0)) : sqrt(fmax(t2h - (2.0 * et * h2 * h2) / den, 0.0));
    }
    t_out[e] = t;
}

struct Workspace {
    void* ptr = nullptr;
    size_t size = 0;
};
Workspace g_ws;

int ensure_ws(size_t bytes) {
    if (bytes <= g_ws.size) return 0;
    if (g_ws.ptr) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(g_ws.ptr));
    }
    g_ws.ptr = nullptr;
    g_ws.size = 0;
    HIP_TRY(hipMalloc(&g_ws.ptr, bytes));
    g_ws.size = bytes;
    return 0;
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// forward with saved activations: Zs[l] (pre-activation) and As[l]
int forward_saved(const float* P, const float* X, int n, float* const* Zs, float* const* As, hipStream_t st) {
    const float* in = X;
    for (int l = 0; l < kLayers; ++l) {
        const long cnt = (long)n * layer_out(l);
        if (l < kLayers - 1)
            launch_fwd<true>(nblk(cnt), st, P, l, in, n, Zs[l], As[l]);
        else
            launch_fwd<false>(nblk(cnt), st, P, l, in, n, Zs[l], As[l]);
        HIP_TRY(hipGetLastError());
        in = As[l];
    }
    return 0;
}

// loss and gradient of the mean squared error; G receives every parameter's gradient
int loss_grad(const float* P, const float* X, const float* T, int n, float* G, float* loss, hipStream_t st) {
    // workspace: Z, A per layer (n x 64 each), dA/dX ping-pong (2 x n x 64), partials
    const size_t per = (size_t)n * kH;
    const size_t need = sizeof(float) * (per * 2 * kLayers + 2 * per + nblk((long)n * kOut) + 16);
    int rc = ensure_ws(need);
    if (rc) return rc;
    float* base = static_cast<float*>(g_ws.ptr);
    float* Zs[kLayers];
    float* As[kLayers];
    for (int l = 0; l < kLayers; ++l) {
        Zs[l] = base + per * (2 * l);
        As[l] = base + per * (2 * l + 1);
    }
    float* d0 = base + per * 2 * kLayers;
    float* d1 = d0 + per;
    float* part = d1 + per;
    rc = forward_saved(P, X, n, Zs, As, st);
    if (rc) return rc;
    const long ny = (long)n * kOut;
    hipLaunchKernelGGL(k_mse, dim3(nblk(ny)), dim3(256), 0, st, As[kLayers - 1], T, n, d0, part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, st, part, (int)nblk(ny), 1.0f / (float)ny, loss);
    HIP_TRY(hipGetLastError());
    float* dA = d0;   // gradient w.r.t. the layer's output (post-activation)
    float* dX = d1;
    for (int l = kLayers - 1; l >= 0; --l) {
        const long cnt = (long)n * layer_out(l);
        if (l < kLayers - 1) {
            hipLaunchKernelGGL(k_silu_bwd, dim3(nblk(cnt)), dim3(256), 0, st, Zs[l], dA, cnt);
            HIP_TRY(hipGetLastError());
        }
        const float* Xin = l == 0 ? X : As[l - 1];
        hipLaunchKernelGGL(k_linear_bwd_w, dim3(layer_out(l) * (layer_in(l) + 1)), dim3(256), 0, st, l, dA, Xin,
                           n, G);
        HIP_TRY(hipGetLastError());
        if (l > 0) {
            if (g_q)
                hipLaunchKernelGGL(k_linear_bwd_x<true>, dim3(nblk((long)n * layer_in(l))), dim3(256), 0, st, P, l,
                                   dA, n, dX);
            else
                hipLaunchKernelGGL(k_linear_bwd_x<false>, dim3(nblk((long)n * layer_in(l))), dim3(256), 0, st, P, l,
                                   dA, n, dX);
            HIP_TRY(hipGetLastError());
            float* tmp = dA;
            dA = dX;
            dX = tmp;
        }
    }
    return 0;
}


// ---------------------------------------------------------------------------
// Fused training: one kernel per step computes the forward pass, the loss and
// the backward pass of a minibatch, a second reduces the per-workgroup
// gradient partials and applies Adam, a third advances the step counter and
// logs the loss.  A run of steps is captured once into a hipGraph and
// replayed, and the minibatch indices are drawn on the device (counter-based
// hash of seed, step and sample), so a training step costs no host work.
//
// A workgroup (256 threads) processes chunks of kS = 32 samples, all of them
// in LDS: the weights (rounded to bf16 in the quantisation-aware steps), the
// inputs, every hidden layer's pre-activation Z and activation A, and two
// gradient buffers.  Each thread owns a 2-sample x 4-output tile of a layer
// in the forward pass and of dA in the backward pass, a 4 x 4 tile of every
// hidden weight gradient (kept in registers over the workgroup's chunks) and
// three entries of the small gradients (layer 0, biases, output layer).
// Rows are padded to kLd floats so that the float4 reads of a wave spread
// over the LDS banks.
// ---------------------------------------------------------------------------
constexpr int kS = 32;                 // samples per chunk
constexpr int kLd = 68;                // padded row stride (floats) of 64-wide rows
constexpr int kLd0 = 8;                // row stride of layer 0 (5 inputs + zero padding)
constexpr int kTrainThreads = 256;
constexpr int kTrainGrid = 256;        // workgroups (at most) of a step
constexpr int kSmall = 3;              // small-gradient entries per thread
// LDS layout (floats)
constexpr int kOffW0 = 0;                              // 64 x kLd0
constexpr int kOffW1 = kOffW0 + kH * kLd0;             // 3 x (64 x kLd)
constexpr int kOffW4 = kOffW1 + 3 * kH * kLd;          // 2 x kLd
constexpr int kOffB = kOffW4 + kOut * kLd;             // 4 x 64 + 2 biases (+pad)
constexpr int kOffX = kOffB + 4 * kH + 8;              // kS x kLd0
constexpr int kOffZ = kOffX + kS * kLd0;               // 4 x (kS x kLd)
constexpr int kOffA = kOffZ + 4 * kS * kLd;            // 4 x (kS x kLd)
constexpr int kOffD0 = kOffA + 4 * kS * kLd;           // kS x kLd
constexpr int kOffD1 = kOffD0 + kS * kLd;              // kS x kLd
constexpr int kOffY = kOffD1 + kS * kLd;               // kS x 2 (dY)
constexpr int kOffL = kOffY + kS * kOut;               // loss partials per wave (4)
constexpr int kTrainLdsFloats = kOffL + 8;
static_assert(kTrainLdsFloats * 4 <= 160 * 1024, "fused training LDS");

struct FusedArgs {
    const float* P;              // parameters (kParams)
    const float* X;              // (n, 5) inputs
    const float* T;              // (n, 2) targets
    long long n;                 // samples in the set
    int batch;                   // samples per step
    int fixed;                   // 1: rows [0, batch) in order (gradient checks); 0: drawn
    const int* step;             // device iteration counter (0-based)
    int steps;                   // iterations of the run (a step >= steps does nothing)
    int q_from;                  // iterations >= q_from use the fused inference's bf16 numerics
    unsigned long long seed;
    float inv_count;             // 1 / (batch * kOut): MSE normalisation
    float* gpart;                // (gridDim.x, kParams) gradient partials
    float* lpart;                // (gridDim.x) loss partials (sum of squared errors)
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    // splitmix64 finaliser (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float silu_f(float z) { return z * sigmoidf_(z); }
__device__ __forceinline__ float silu_d(float z) {
    const float sg = sigmoidf_(z);
    return sg * (1.0f + z * (1.0f - sg));
}

// small-gradient entry e: layer 0 (64 x (5 + bias)), hidden biases (3 x 64),
// output layer (2 x (64 + bias)); returns the parameter index and sets kind,
// layer, o, i (i == fan-in: the bias)
__device__ __forceinline__ int small_entry(int e, int& kind, int& l, int& o, int& i) {
    if (e < kH * (kIn + 1)) {
        kind = 0; l = 0; o = e / (kIn + 1); i = e % (kIn + 1);
        return i < kIn ? layer_w_off(0) + o * kIn + i : layer_b_off(0) + o;
    }
    e -= kH * (kIn + 1);
    if (e < 3 * kH) {
        kind = 1; l = 1 + e / kH; o = e % kH; i = kH;
        return layer_b_off(l) + o;
    }
    e -= 3 * kH;
    if (e < kOut * (kH + 1)) {
        kind = 2; l = 4; o = e / (kH + 1); i = e % (kH + 1);
        return i < kH ? layer_w_off(4) + o * kH + i : layer_b_off(4) + o;
    }
    kind = -1; l = o = i = 0;
    return -1;
}

template <bool Q>
__device__ void train_body(const FusedArgs& a, float* sm, int it) {
    const int tid = threadIdx.x;
    float* W0 = sm + kOffW0;
    float* Bs = sm + kOffB;
    float* Xs = sm + kOffX;
    float* dY = sm + kOffY;
    auto Wl = [&](int l) -> float* { return l == 4 ? sm + kOffW4 : sm + kOffW1 + (l - 1) * kH * kLd; };
    auto Zl = [&](int l) -> float* { return sm + kOffZ + l * kS * kLd; };
    auto Al = [&](int l) -> float* { return sm + kOffA + l * kS * kLd; };
    auto bias = [&](int l) -> const float* { return Bs + l * kH; };

    // weights (rounded to bf16 in the quantisation-aware numerics: the forward
    // and the backward through the inputs use the rounded weights) and biases
    for (int e = tid; e < kH * kLd0; e += kTrainThreads) {
        const int o = e / kLd0, i = e % kLd0;
        const float w = i < kIn ? a.P[layer_w_off(0) + o * kIn + i] : 0.0f;
        W0[e] = Q ? bf16r(w) : w;
    }
    for (int l = 1; l <= 4; ++l) {
        const int out = layer_out(l);
        for (int e = tid; e < out * kLd; e += kTrainThreads) {
            const int o = e / kLd, i = e % kLd;
            const float w = i < kH ? a.P[layer_w_off(l) + o * kH + i] : 0.0f;
            Wl(l)[e] = Q ? bf16r(w) : w;
        }
    }
    for (int e = tid; e < 4 * kH + kOut; e += kTrainThreads) {
        const int l = e / kH, o = e % kH;
        Bs[e] = l < 4 ? a.P[layer_b_off(l) + o] : a.P[layer_b_off(4) + (e - 4 * kH)];
    }

    // gradient accumulators
    const int fs = (tid >> 4) * 2;      // forward / dA tiles: samples fs, fs + 1
    const int fo = (tid & 15) * 4;      //                    outputs (or inputs) fo .. fo + 3
    const int go = (tid >> 4) * 4;      // weight-gradient tiles: outputs go .. go + 3
    const int gi = (tid & 15) * 4;      //                        inputs gi .. gi + 3
    float gw[3][4][4];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) gw[l][r][c] = 0.0f;
    float gs[kSmall] = {0.0f, 0.0f, 0.0f};
    int sk[kSmall], sl[kSmall], so[kSmall], si[kSmall], sp[kSmall];
#pragma unroll
    for (int j = 0; j < kSmall; ++j) sp[j] = small_entry(tid + j * kTrainThreads, sk[j], sl[j], so[j], si[j]);
    float lsum = 0.0f;

    const int n_chunks = (a.batch + kS - 1) / kS;
    for (int ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
        __syncthreads();  // the weights are loaded / the previous chunk is done
        // inputs and targets of the chunk's samples (rows past the batch: zero
        // input, zero loss gradient)
        if (tid < kS) {
            const int g = ch * kS + tid;
            const bool valid = g < a.batch;
            long long row = 0;
            if (valid) row = a.fixed ? g : (long long)(mix64(a.seed ^ mix64(((unsigned long long)it << 32) ^ (unsigned)g)) %
                                                       (unsigned long long)a.n);
            for (int i = 0; i < kLd0; ++i) {
                float xv = (valid && i < kIn) ? a.X[row * kIn + i] : 0.0f;
                if (Q) xv = bf16pair(xv);
                Xs[tid * kLd0 + i] = xv;
            }
            dY[tid * kOut] = valid ? a.T[row * kOut] : 0.0f;       // targets for now
            dY[tid * kOut + 1] = valid ? a.T[row * kOut + 1] : 0.0f;
            Xs[tid * kLd0 + 7] = valid ? 1.0f : 0.0f;             // validity flag in the padding
        }
        __syncthreads();

        // forward: hidden layers
        for (int l = 0; l < 4; ++l) {
            const float* in = l == 0 ? Xs : Al(l - 1);
            const int ld_in = l == 0 ? kLd0 : kLd;
            const float* W = l == 0 ? W0 : Wl(l);
            const int ld_w = l == 0 ? kLd0 : kLd;
            const int n_in = l == 0 ? kLd0 : kH;   // layer 0: padded inputs times zero weights
            float z[2][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) z[0][c] = z[1][c] = bias(l)[fo + c];
            for (int i = 0; i < n_in; i += 4) {
                float4 a0 = *reinterpret_cast<const float4*>(in + fs * ld_in + i);
                float4 a1 = *reinterpret_cast<const float4*>(in + (fs + 1) * ld_in + i);
                if (l == 0 && i == 4) { a0.w = 0.0f; a1.w = 0.0f; }  // the validity flag is no input
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float4 w = *reinterpret_cast<const float4*>(W + (fo + c) * ld_w + i);
                    z[0][c] = fmaf(a0.x, w.x, fmaf(a0.y, w.y, fmaf(a0.z, w.z, fmaf(a0.w, w.w, z[0][c]))));
                    z[1][c] = fmaf(a1.x, w.x, fmaf(a1.y, w.y, fmaf(a1.z, w.z, fmaf(a1.w, w.w, z[1][c]))));
                }
            }
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                float4 zz = make_float4(z[r][0], z[r][1], z[r][2], z[r][3]);
                float4 aa = make_float4(silu_f(zz.x), silu_f(zz.y), silu_f(zz.z), silu_f(zz.w));
                if (Q) aa = make_float4(bf16pair(aa.x), bf16pair(aa.y), bf16pair(aa.z), bf16pair(aa.w));
                *reinterpret_cast<float4*>(Zl(l) + (fs + r) * kLd + fo) = zz;
                *reinterpret_cast<float4*>(Al(l) + (fs + r) * kLd + fo) = aa;
            }
            __syncthreads();
        }
        // output layer, loss and dY = 2 (y - t) / (batch kOut)
        if (tid < kS * kOut) {
            const int s = tid >> 1, o = tid & 1;
            const float* in = Al(3) + s * kLd;
            const float* W = Wl(4) + o * kLd;
            float y = Bs[4 * kH + o];
            for (int i = 0; i < kH; i += 4) {
                const float4 av = *reinterpret_cast<const float4*>(in + i);
                const float4 w = *reinterpret_cast<const float4*>(W + i);
                y = fmaf(av.x, w.x, fmaf(av.y, w.y, fmaf(av.z, w.z, fmaf(av.w, w.w, y))));
            }
            const bool valid = Xs[s * kLd0 + 7] != 0.0f;
            const float d = valid ? y - dY[s * kOut + o] : 0.0f;
            lsum += d * d;
            dY[s * kOut + o] = 2.0f * d * a.inv_count;
        }
        __syncthreads();
        // output-layer gradients (small entries) and dA3 = dY W4
#pragma unroll
        for (int j = 0; j < kSmall; ++j) {
            if (sk[j] != 2) continue;
            float acc = 0.0f;
            for (int s = 0; s < kS; ++s)
                acc = fmaf(dY[s * kOut + so[j]], si[j] < kH ? Al(3)[s * kLd + si[j]] : 1.0f, acc);
            gs[j] += acc;
        }
        float* dA = sm + kOffD0;
        float* dB = sm + kOffD1;
        {
            const float* W = Wl(4);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                float4 v;
                const float y0 = dY[(fs + r) * kOut], y1 = dY[(fs + r) * kOut + 1];
                v.x = fmaf(y0, W[fo], y1 * W[kLd + fo]);
                v.y = fmaf(y0, W[fo + 1], y1 * W[kLd + fo + 1]);
                v.z = fmaf(y0, W[fo + 2], y1 * W[kLd + fo + 2]);
                v.w = fmaf(y0, W[fo + 3], y1 * W[kLd + fo + 3]);
                *reinterpret_cast<float4*>(dA + (fs + r) * kLd + fo) = v;
            }
        }
        __syncthreads();
        // hidden layers 3 .. 0
        for (int l = 3; l >= 0; --l) {
            // dZ = dA * silu'(Z), in place
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                float4 d = *reinterpret_cast<float4*>(dA + (fs + r) * kLd + fo);
                const float4 z = *reinterpret_cast<const float4*>(Zl(l) + (fs + r) * kLd + fo);
                d.x *= silu_d(z.x);
                d.y *= silu_d(z.y);
                d.z *= silu_d(z.z);
                d.w *= silu_d(z.w);
                *reinterpret_cast<float4*>(dA + (fs + r) * kLd + fo) = d;
            }
            __syncthreads();
            if (l > 0) {
                // weight gradient tile: sum over the chunk's samples of dZ[s][o] A[l-1][s][i]
                const float* Ap = Al(l - 1);
                float acc[4][4];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0f;
                for (int s = 0; s < kS; ++s) {
                    const float4 dz = *reinterpret_cast<const float4*>(dA + s * kLd + go);
                    const float4 av = *reinterpret_cast<const float4*>(Ap + s * kLd + gi);
                    const float dzv[4] = {dz.x, dz.y, dz.z, dz.w};
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        acc[r][0] = fmaf(dzv[r], av.x, acc[r][0]);
                        acc[r][1] = fmaf(dzv[r], av.y, acc[r][1]);
                        acc[r][2] = fmaf(dzv[r], av.z, acc[r][2]);
                        acc[r][3] = fmaf(dzv[r], av.w, acc[r][3]);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) gw[l - 1][r][c] += acc[r][c];
                // dA[l-1] = dZ W_l (the weights as the forward used them)
                const float* W = Wl(l);
                float v[2][4];
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) v[r][c] = 0.0f;
                for (int o = 0; o < kH; o += 4) {
                    const float4 d0 = *reinterpret_cast<const float4*>(dA + fs * kLd + o);
                    const float4 d1 = *reinterpret_cast<const float4*>(dA + (fs + 1) * kLd + o);
                    const float dd0[4] = {d0.x, d0.y, d0.z, d0.w};
                    const float dd1[4] = {d1.x, d1.y, d1.z, d1.w};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float4 w = *reinterpret_cast<const float4*>(W + (o + q) * kLd + fo);
                        v[0][0] = fmaf(dd0[q], w.x, v[0][0]);
                        v[0][1] = fmaf(dd0[q], w.y, v[0][1]);
                        v[0][2] = fmaf(dd0[q], w.z, v[0][2]);
                        v[0][3] = fmaf(dd0[q], w.w, v[0][3]);
                        v[1][0] = fmaf(dd1[q], w.x, v[1][0]);
                        v[1][1] = fmaf(dd1[q], w.y, v[1][1]);
                        v[1][2] = fmaf(dd1[q], w.z, v[1][2]);
                        v[1][3] = fmaf(dd1[q], w.w, v[1][3]);
                    }
                }
#pragma unroll
                for (int r = 0; r < 2; ++r)
                    *reinterpret_cast<float4*>(dB + (fs + r) * kLd + fo) = make_float4(v[r][0], v[r][1], v[r][2], v[r][3]);
            }
            // bias (and layer-0) gradients from dZ of this layer
#pragma unroll
            for (int j = 0; j < kSmall; ++j) {
                if (!((sk[j] == 1 && sl[j] == l) || (sk[j] == 0 && l == 0))) continue;
                float acc = 0.0f;
                for (int s = 0; s < kS; ++s) {
                    const float xin = (sk[j] == 0 && si[j] < kIn) ? Xs[s * kLd0 + si[j]] : 1.0f;
                    acc = fmaf(dA[s * kLd + so[j]], xin, acc);
                }
                gs[j] += acc;
            }
            __syncthreads();
            float* t = dA;
            dA = dB;
            dB = t;
        }
    }
    // this workgroup's partials: every parameter written once
    float* gp = a.gpart + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int l = 1; l <= 3; ++l)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            *reinterpret_cast<float4*>(gp + layer_w_off(l) + (go + r) * kH + gi) =
                make_float4(gw[l - 1][r][0], gw[l - 1][r][1], gw[l - 1][r][2], gw[l - 1][r][3]);
#pragma unroll
    for (int j = 0; j < kSmall; ++j)
        if (sp[j] >= 0) gp[sp[j]] = gs[j];
    // loss partial
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
    float* lw = sm + kOffL;
    if ((tid & 63) == 0) lw[tid >> 6] = lsum;
    __syncthreads();
    if (tid == 0) a.lpart[blockIdx.x] = lw[0] + lw[1] + lw[2] + lw[3];
}

__global__ __launch_bounds__(kTrainThreads) void k_train_fused(FusedArgs a) {
    extern __shared__ __align__(16) float tsm[];
    const int it = __builtin_amdgcn_readfirstlane(a.step ? *a.step : 0);
    if (it >= a.steps) return;
    if (it >= a.q_from) train_body<true>(a, tsm, it);
    else train_body<false>(a, tsm, it);
}

// gradient = sum of the workgroup partials; then Adam (Kingma & Ba 2015) with
// bias correction and the cosine-decayed learning rate of iteration `it`
__global__ void k_train_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                             const float* __restrict__ gpart, int n_part, const int* __restrict__ step, int steps,
                             int step0, float lr, float lr_final, float b1, float b2, float eps) {
    const int it = *step;
    if (it >= steps) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kParams) return;
    float g = 0.0f;
    for (int w = 0; w < n_part; ++w) g += gpart[(size_t)w * kParams + e];
    const float c = cosf(3.14159265358979f * (float)it / (float)max(steps - 1, 1));
    const float a = lr_final + 0.5f * (lr - lr_final) * (1.0f + c);
    const int t = step0 + it + 1;
    const float bc1 = 1.0f - powf(b1, (float)t), bc2 = 1.0f - powf(b2, (float)t);
    const float m = b1 * M[e] + (1.0f - b1) * g;
    const float v = b2 * V[e] + (1.0f - b2) * g * g;
    M[e] = m;
    V[e] = v;
    P[e] -= a * (m / bc1) / (sqrtf(v / bc2) + eps);
}

// step bookkeeping: the loss of logged iterations, then the counter advances
__global__ void k_train_advance(int* step, int steps, const float* __restrict__ lpart, int n_part, float inv_count,
                                int log_every, float* __restrict__ loss_log) {
    const int it = *step;
    if (it >= steps) return;
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n_part; i += blockDim.x) s += lpart[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (loss_log && log_every > 0 && (it % log_every == 0 || it == steps - 1))
            loss_log[it / log_every + (it % log_every != 0 ? 1 : 0)] = red[0] * inv_count;
        *step = it + 1;
    }
}

__global__ void k_set_step(int* step, int v) { *step = v; }

__global__ void k_reduce_grads(const float* __restrict__ gpart, int n_part, const float* __restrict__ lpart,
                               float inv_count, float* __restrict__ G, float* __restrict__ loss) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kParams) {
        float g = 0.0f;
        for (int w = 0; w < n_part; ++w) g += gpart[(size_t)w * kParams + e];
        G[e] = g;
    }
    if (e == 0) {
        float s = 0.0f;
        for (int w = 0; w < n_part; ++w) s += lpart[w];
        *loss = s * inv_count;
    }
}

int set_fused_lds() {
    static bool done = false;
    if (done) return 0;
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_train_fused),
                                hipFuncAttributeMaxDynamicSharedMemorySize, kTrainLdsFloats * 4));
    done = true;
    return 0;
}

}  // namespace

extern "C" {

int pstm_mlp_param_count(void) { return kParams; }

int pstm_mlp_set_precision(int bf16) {
    g_err.clear();
    if (bf16 != 0 && bf16 != 1) return fail(1, "pstm_mlp_set_precision: expected 0 (f32) or 1 (bf16)");
    g_q = bf16 == 1;
    return 0;
}

const char* pstm_mlp_last_error(void) { return g_err.c_str(); }

int pstm_mlp_forward(const float* params, const float* x, int n, float* y, void* stream) {
    g_err.clear();
    if (!params || !x || !y || n < 0) return fail(1, "pstm_mlp_forward: bad arguments");
    if (n == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t per = (size_t)n * kH;
    int rc = ensure_ws(sizeof(float) * per * 2);
    if (rc) return rc;
    float* a0 = static_cast<float*>(g_ws.ptr);
    float* a1 = a0 + per;
    const float* in = x;
    for (int l = 0; l < kLayers; ++l) {
        float* out = l == kLayers - 1 ? y : (l % 2 == 0 ? a0 : a1);
        const long cnt = (long)n * layer_out(l);
        if (l < kLayers - 1)
            launch_fwd<true>(nblk(cnt), st, params, l, in, n, nullptr, out);
        else
            launch_fwd<false>(nblk(cnt), st, params, l, in, n, nullptr, out);
        HIP_TRY(hipGetLastError());
        in = out;
    }
    return 0;
}

int pstm_mlp_loss_grad(const float* params, const float* x, const float* target, int n, float* grads,
                       float* loss, void* stream) {
    g_err.clear();
    if (!params || !x || !target || !grads || !loss || n <= 0) return fail(1, "pstm_mlp_loss_grad: bad arguments");
    return loss_grad(params, x, target, n, grads, loss, static_cast<hipStream_t>(stream));
}

int pstm_mlp_adam(float* params, const float* grads, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int step, void* stream) {
    g_err.clear();
    if (!params || !grads || !m || !v || step < 1) return fail(1, "pstm_mlp_adam: bad arguments");
    const float bc1 = 1.0f - powf(beta1, (float)step);
    const float bc2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(kParams)), dim3(256), 0, static_cast<hipStream_t>(stream), params, grads,
                       m, v, kParams, lr, beta1, beta2, eps, bc1, bc2);
    HIP_TRY(hipGetLastError());
    return 0;
}

int pstm_mlp_train_step(float* params, float* m, float* v, float* grads, const float* x, const float* target,
                        int n, float lr, float beta1, float beta2, float eps, int step, float* loss,
                        void* stream) {
    int rc = pstm_mlp_loss_grad(params, x, target, n, grads, loss, stream);
    if (rc) return rc;
    return pstm_mlp_adam(params, grads, m, v, lr, beta1, beta2, eps, step, stream);
}

// Fused loss and gradient over rows [0, n) of (x, target), in the current
// precision: the gradient check of the fused training kernel.
int pstm_mlp_loss_grad_fused(const float* params, const float* x, const float* target, int n, float* grads,
                             float* loss, void* stream) {
    g_err.clear();
    if (!params || !x || !target || !grads || !loss || n <= 0)
        return fail(1, "pstm_mlp_loss_grad_fused: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = set_fused_lds();
    if (rc) return rc;
    const int n_chunks = (n + kS - 1) / kS;
    const int grid = std::min(kTrainGrid, n_chunks);
    rc = ensure_ws(sizeof(float) * ((size_t)grid * kParams + grid + 64));
    if (rc) return rc;
    float* gpart = static_cast<float*>(g_ws.ptr);
    float* lpart = gpart + (size_t)grid * kParams;
    FusedArgs fa{};
    fa.P = params;
    fa.X = x;
    fa.T = target;
    fa.n = n;
    fa.batch = n;
    fa.fixed = 1;
    fa.step = nullptr;
    fa.steps = 1;
    fa.q_from = g_q ? 0 : INT_MAX;
    fa.inv_count = 1.0f / ((float)n * kOut);
    fa.gpart = gpart;
    fa.lpart = lpart;
    hipLaunchKernelGGL(k_train_fused, dim3(grid), dim3(kTrainThreads), kTrainLdsFloats * 4, st, fa);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_grads, dim3(nblk(kParams)), dim3(256), 0, st, gpart, grid, lpart, fa.inv_count,
                       grads, loss);
    HIP_TRY(hipGetLastError());
    return 0;
}

int pstm_mlp_train(float* params, float* m, float* v, const float* x, const float* target, long long n, int steps,
                   int batch, float lr, float lr_final, float beta1, float beta2, float eps, int step0, int q_from,
                   unsigned long long seed, int log_every, float* loss_log, void* stream) {
    g_err.clear();
    if (!params || !m || !v || !x || !target || n <= 0 || steps < 0 || batch <= 0 || step0 < 0 || log_every < 0)
        return fail(1, "pstm_mlp_train: bad arguments");
    if (steps == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = set_fused_lds();
    if (rc) return rc;
    const int n_chunks = (batch + kS - 1) / kS;
    const int grid = std::min(kTrainGrid, n_chunks);
    const size_t need = sizeof(float) * ((size_t)grid * kParams + grid + 64);
    void* mem = nullptr;
    HIP_TRY(hipMalloc(&mem, need));
    float* gpart = static_cast<float*>(mem);
    float* lpart = gpart + (size_t)grid * kParams;
    int* step = reinterpret_cast<int*>(lpart + grid + 32);
    FusedArgs fa{};
    fa.P = params;
    fa.X = x;
    fa.T = target;
    fa.n = n;
    fa.batch = batch;
    fa.fixed = 0;
    fa.step = step;
    fa.steps = steps;
    fa.q_from = q_from;
    fa.seed = seed;
    fa.inv_count = 1.0f / ((float)batch * kOut);
    fa.gpart = gpart;
    fa.lpart = lpart;
    // the iterations of one replay: kGraphSteps x (fused step, Adam, advance);
    // replays past `steps` do nothing (every kernel tests the counter)
    constexpr int kGraphSteps = 64;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        (void)hipFree(mem);
    };
    auto enqueue = [&](hipStream_t s) -> hipError_t {
        for (int k = 0; k < kGraphSteps; ++k) {
            hipLaunchKernelGGL(k_train_fused, dim3(grid), dim3(kTrainThreads), kTrainLdsFloats * 4, s, fa);
            hipLaunchKernelGGL(k_train_adam, dim3(nblk(kParams)), dim3(256), 0, s, params, m, v, gpart, grid, step,
                               steps, step0, lr, lr_final, beta1, beta2, eps);
            hipLaunchKernelGGL(k_train_advance, dim3(1), dim3(256), 0, s, step, steps, lpart, grid, fa.inv_count,
                               log_every, loss_log);
        }
        return hipGetLastError();
    };
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, st, step, 0);
        e = hipGetLastError();
    }
    // the capture stream starts after the caller's stream work (counter reset)
    hipEvent_t ready = nullptr;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ready, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (ready) (void)hipEventDestroy(ready);
    if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t e1 = enqueue(cap);
        e = hipStreamEndCapture(cap, &graph);
        if (e == hipSuccess) e = e1;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    const int replays = (steps + kGraphSteps - 1) / kGraphSteps;
    for (int r = 0; r < replays && e == hipSuccess; ++r) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return fail(3, std::string("pstm_mlp_train: ") + hipGetErrorString(e));
    return 0;
}

int pstm_vti_leg_times(const double* tau, const double* h, const double* v, const double* eta, long long n,
                       double* t_out, void* stream) {
    g_err.clear();
    if (n < 0 || (n > 0 && (!tau || !h || !v || !eta || !t_out))) return fail(1, "pstm_vti_leg_times: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_vti_targets, dim3(nblk(n)), dim3(256), 0, static_cast<hipStream_t>(stream), tau, h, v,
                       eta, (long)n, t_out);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
