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
// two MFMA products per layer), f32 accumulation and biases.  Training with
// Q = true is quantisation-aware (straight-through gradients).
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
