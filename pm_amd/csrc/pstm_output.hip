// pstm_output.hip — the executor's accumulation and output step on the device.
//
// After each tile the reference adds per-output-bin trace-header sums (trace
// count, offset, azimuth sine and cosine) for every trace of the tile's
// aperture query (pstm/pipeline/executor.py:1783-1806); in gather mode each
// offset / offset-vector-tile bin keeps only the traces that pass its filter
// (executor.py:1537-1554) and, at finalize, its image is divided by its fold
// (executor.py:2155-2160).  The traces, images and folds already live in HBM,
// so these steps run there as three small kernels:
//
//   * k_bin_headers_lds — the tile's bins summed in LDS per workgroup, then
//     added to the (nx, ny) header arrays with hardware atomics (k_bin_headers,
//     one global atomic per row, for tiles too large for LDS).  The bin index
//     and the offset are computed with the reference's operations in f64 with
//     explicit rounding (no fused multiply-add), so the bin of every trace is
//     the reference's bin;
//   * k_select_rows — the gather-bin filter as a byte mask per row (the host
//     compacts it in row order);
//   * k_fold_normalize — HBM-bound elementwise pass, 16-byte accesses.
//
// Built with -Xclang -target-feature -Xclang -packed-fp32-ops.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pstm_hip.h"

void pstm_set_last_error(const std::string& msg);

namespace {

int fail(int code, const std::string& msg) {
    pstm_set_last_error(msg);
    return code;
}

#define OUT_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return fail(3, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

constexpr int kBlock = 256;
constexpr size_t kHeaderLds = 160 * 1024;   // LDS bins of one tile: up to 5851 bins

// offset = sqrt(dx^2 + dy^2), rounded as NumPy rounds it (compute_offset,
// pstm/utils/units.py:180-200).  Contraction is switched off wherever the
// reference's rounding matters: the __d*_rn intrinsics alone do not keep the
// compiler from fusing a product and a sum into one FMA.
__device__ __forceinline__ double offset_of(double dx, double dy) {
#pragma clang fp contract(off)
    return __dsqrt_rn(dx * dx + dy * dy);
}

struct BinTerms {
    int bin;        // tile-local bin, ix * tny + iy
    double off, s, c;
};

__device__ __forceinline__ BinTerms bin_terms(const double* __restrict__ sx, const double* __restrict__ sy,
                                              const double* __restrict__ rx, const double* __restrict__ ry, int r,
                                              double x_min, double y_min, double dx, double dy, int tnx, int tny) {
    const double ax = sx[r], ay = sy[r], bx = rx[r], by = ry[r];
    // midpoint (sx + rx) / 2 (offset_to_midpoint, units.py:157-177); halving is exact
    const double mx = __dadd_rn(ax, bx) * 0.5;
    const double my = __dadd_rn(ay, by) * 0.5;
    // tile-local bin, floor((m - min) / d) clipped to the tile (executor.py:1796-1802)
    const double fx = floor(__ddiv_rn(__dsub_rn(mx, x_min), dx));
    const double fy = floor(__ddiv_rn(__dsub_rn(my, y_min), dy));
    // clip in f64 first: a far trace's quotient can exceed the int range
    const int ix = static_cast<int>(fmin(fmax(fx, 0.0), static_cast<double>(tnx - 1)));
    const int iy = static_cast<int>(fmin(fmax(fy, 0.0), static_cast<double>(tny - 1)));
    const double ox = __dsub_rn(bx, ax), oy = __dsub_rn(by, ay);
    BinTerms t{ix * tny + iy, offset_of(ox, oy), 0.0, 1.0};
    // azimuth a = atan2(ox, oy) (north = 0): sin a = ox / |o|, cos a = oy / |o|;
    // a zero offset has a = 0
    if (t.off > 0.0) {
        t.s = __ddiv_rn(ox, t.off);
        t.c = __ddiv_rn(oy, t.off);
    }
    return t;
}

// Most rows of a tile's query lie in its aperture halo and clip into a few
// edge bins, so global atomics would queue on a handful of addresses.  Each
// workgroup therefore sums its share of the rows in an LDS copy of the
// tile's bins (count i32, then offset / sin / cos f64) and adds only the
// bins it touched to the global arrays.
constexpr int kHBlock = 1024;

__global__ void __launch_bounds__(kHBlock)
k_bin_headers_lds(const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ rx,
                  const double* __restrict__ ry, const int32_t* __restrict__ rows, int n, int per_block,
                  double x_min, double y_min, double dx, double dy, int x_start, int y_start, int tnx, int tny,
                  int grid_ny, int* __restrict__ count, double* __restrict__ osum, double* __restrict__ ssum,
                  double* __restrict__ csum) {
    extern __shared__ double lds[];
    const int nb = tnx * tny;
    double* l_off = lds;
    double* l_s = lds + nb;
    double* l_c = lds + 2 * nb;
    int* l_n = reinterpret_cast<int*>(lds + 3 * nb);
    for (int b = threadIdx.x; b < nb; b += kHBlock) {
        l_off[b] = 0.0;
        l_s[b] = 0.0;
        l_c[b] = 0.0;
        l_n[b] = 0;
    }
    __syncthreads();
    const int i0 = blockIdx.x * per_block;
    const int i1 = static_cast<int>(min(static_cast<long long>(n), static_cast<long long>(i0) + per_block));
    for (int i = i0 + static_cast<int>(threadIdx.x); i < i1; i += kHBlock) {
        const BinTerms t = bin_terms(sx, sy, rx, ry, rows ? rows[i] : i, x_min, y_min, dx, dy, tnx, tny);
        atomicAdd(l_n + t.bin, 1);
        atomicAdd(l_off + t.bin, t.off);
        atomicAdd(l_s + t.bin, t.s);
        atomicAdd(l_c + t.bin, t.c);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kHBlock) {
        const int k = l_n[b];
        if (k == 0) continue;
        const long long cell = static_cast<long long>(b / tny + x_start) * grid_ny + (b % tny + y_start);
        atomicAdd(count + cell, k);
        unsafeAtomicAdd(osum + cell, l_off[b]);
        unsafeAtomicAdd(ssum + cell, l_s[b]);
        unsafeAtomicAdd(csum + cell, l_c[b]);
    }
}

// tiles whose bins do not fit in LDS: one global atomic per row and sum
__global__ void __launch_bounds__(kBlock)
k_bin_headers(const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ rx,
              const double* __restrict__ ry, const int32_t* __restrict__ rows, int n, double x_min,
              double y_min, double dx, double dy, int x_start, int y_start, int tnx, int tny, int grid_ny,
              int* __restrict__ count, double* __restrict__ osum, double* __restrict__ ssum,
              double* __restrict__ csum) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const BinTerms t = bin_terms(sx, sy, rx, ry, rows ? rows[i] : i, x_min, y_min, dx, dy, tnx, tny);
    const long long cell = static_cast<long long>(t.bin / tny + x_start) * grid_ny + (t.bin % tny + y_start);
    atomicAdd(count + cell, 1);
    unsafeAtomicAdd(osum + cell, t.off);
    unsafeAtomicAdd(ssum + cell, t.s);
    unsafeAtomicAdd(csum + cell, t.c);
}

__global__ void __launch_bounds__(kBlock)
k_select_rows(const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ rx,
              const double* __restrict__ ry, const int32_t* __restrict__ rows, int n, int checks,
              double off_lo, double off_hi, double ox_lo, double ox_hi, double oy_lo, double oy_hi,
              uint8_t* __restrict__ keep) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int r = rows ? rows[i] : i;
    const double ox = __dsub_rn(rx[r], sx[r]), oy = __dsub_rn(ry[r], sy[r]);
    bool k = true;
    if (checks & (PSTM_SEL_OFFSET_MIN | PSTM_SEL_OFFSET_MAX)) {
        const double off = offset_of(ox, oy);
        if (checks & PSTM_SEL_OFFSET_MIN) k = k && (off >= off_lo);
        if (checks & PSTM_SEL_OFFSET_MAX) k = k && (off <= off_hi);
    }
    if (checks & PSTM_SEL_OVT_X_MIN) k = k && (ox >= ox_lo);
    if (checks & PSTM_SEL_OVT_X_MAX) k = k && (ox <= ox_hi);
    if (checks & PSTM_SEL_OVT_Y_MIN) k = k && (oy >= oy_lo);
    if (checks & PSTM_SEL_OVT_Y_MAX) k = k && (oy <= oy_hi);
    keep[i] = k ? 1 : 0;
}

// out[p, k] = image[p, k] / fold[p] where fold[p] > 0, else 0; four samples
// per thread (nt % 4 == 0 and 16-byte aligned pointers), grid-stride
__global__ void __launch_bounds__(kBlock)
k_fold_normalize4(const float4* image, const int32_t* __restrict__ fold, long long n4, int nt4, float4* out) {
    for (long long e = blockIdx.x * (long long)kBlock + threadIdx.x; e < n4; e += (long long)gridDim.x * kBlock) {
        const int f = fold[e / nt4];
        float4 v = image[e];
        if (f > 0) {
            const float d = static_cast<float>(f);
            v.x = v.x / d;
            v.y = v.y / d;
            v.z = v.z / d;
            v.w = v.w / d;
        } else {
            v = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        out[e] = v;
    }
}

__global__ void __launch_bounds__(kBlock)
k_fold_normalize1(const float* image, const int32_t* __restrict__ fold, long long n, int nt, float* out) {
    for (long long e = blockIdx.x * (long long)kBlock + threadIdx.x; e < n; e += (long long)gridDim.x * kBlock) {
        const int f = fold[e / nt];
        out[e] = f > 0 ? image[e] / static_cast<float>(f) : 0.f;
    }
}

// Time-variant image back onto the uniform axis: out[p, k] is np.interp of
// x_out[k] over (xp, image[p, :]) with numpy's arithmetic in f64 — the
// bracket j is the last xp[j] <= x; below xp[0] the first value, above the
// last xp the last value, on a node its value, else
// slope * (x - xp[j]) + fp[j] with slope = (fp[j+1] - fp[j]) / (xp[j+1] - xp[j])
// (and the form from xp[j+1] if that is NaN) — rounded to f32.  The bracket
// is searched once per output sample and reused for kRsPil pillars; the
// workgroup first copies xp to LDS (up to kRsLds nodes), so that the
// bisection's chain of dependent reads stays out of global memory.
constexpr int kRsPil = 16;
constexpr int kRsLds = 4096;

__global__ void __launch_bounds__(kBlock)
k_resample_uniform(const float* __restrict__ tv, const double* __restrict__ xp_g, int n_tv,
                   const double* __restrict__ x_out, int n_out, long long n_pillars, float* __restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ double xs[];  // n_tv doubles when in_lds (the launch sizes it)
    const bool in_lds = n_tv <= kRsLds;
    if (in_lds)
        for (int i = threadIdx.x; i < n_tv; i += kBlock) xs[i] = xp_g[i];
    __syncthreads();
    const double* xp = in_lds ? xs : xp_g;
    const int k = blockIdx.y * kBlock + threadIdx.x;
    if (k >= n_out) return;
    const long long p0 = static_cast<long long>(blockIdx.x) * kRsPil;
    const long long p1 = p0 + kRsPil < n_pillars ? p0 + kRsPil : n_pillars;
    if (n_tv < 2) {  // the reference leaves the uniform image zero
        for (long long p = p0; p < p1; ++p) out[p * n_out + k] = 0.f;
        return;
    }
    const double x = x_out[k];
    // 0: interpolate at j; 1: value at j
    int mode = 1, j;
    if (x > xp[n_tv - 1]) {
        j = n_tv - 1;
    } else if (x < xp[0]) {
        j = 0;
    } else {
        int lo = 0, hi = n_tv;  // first index with xp > x, by bisection
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (x >= xp[mid]) lo = mid + 1;
            else hi = mid;
        }
        j = lo - 1;
        mode = (j == n_tv - 1 || xp[j] == x) ? 1 : 0;
    }
    const double x0 = xp[j];
    const double x1 = mode ? x0 : xp[j + 1];
#pragma unroll 4
    for (long long p = p0; p < p1; ++p) {
        const float* fp = tv + p * n_tv;
        double r;
        if (mode) {
            r = fp[j];
        } else {
            const double f0 = fp[j], f1 = fp[j + 1];
            const double slope = (f1 - f0) / (x1 - x0);
            r = slope * (x - x0) + f0;
            if (isnan(r)) {
                r = slope * (x - x1) + f1;
                if (isnan(r) && f0 == f1) r = f0;
            }
        }
        out[p * n_out + k] = static_cast<float>(r);
    }
}

int grid_for(long long work) {
    long long b = (work + kBlock - 1) / kBlock;
    if (b > 16384) b = 16384;
    return static_cast<int>(b < 1 ? 1 : b);
}

}  // namespace

extern "C" {

int pstm_hip_bin_headers(const double* source_x, const double* source_y, const double* receiver_x,
                         const double* receiver_y, const int32_t* trace_index, int32_t n_traces, double x_min,
                         double y_min, double dx, double dy, int32_t x_start, int32_t y_start, int32_t tile_nx,
                         int32_t tile_ny, int32_t grid_nx, int32_t grid_ny, int32_t* trace_count,
                         double* offset_sum, double* azimuth_sin_sum, double* azimuth_cos_sum, void* stream) {
    pstm_set_last_error("");
    if (n_traces < 0) return fail(1, "pstm_hip_bin_headers: n_traces < 0");
    if (tile_nx < 1 || tile_ny < 1 || x_start < 0 || y_start < 0 || x_start + tile_nx > grid_nx ||
        y_start + tile_ny > grid_ny)
        return fail(1, "pstm_hip_bin_headers: tile [" + std::to_string(x_start) + ", " +
                           std::to_string(x_start + tile_nx) + ") x [" + std::to_string(y_start) + ", " +
                           std::to_string(y_start + tile_ny) + ") outside the " + std::to_string(grid_nx) +
                           " x " + std::to_string(grid_ny) + " grid");
    if (!(dx > 0.0) || !(dy > 0.0)) return fail(1, "pstm_hip_bin_headers: bin sizes must be > 0");
    if (n_traces == 0) return 0;
    if (!source_x || !source_y || !receiver_x || !receiver_y || !trace_count || !offset_sum || !azimuth_sin_sum ||
        !azimuth_cos_sum)
        return fail(1, "pstm_hip_bin_headers: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const long long nb = static_cast<long long>(tile_nx) * tile_ny;
    const size_t lds = static_cast<size_t>(nb) * (3 * sizeof(double) + sizeof(int));
    if (lds <= kHeaderLds) {
        static bool attr_set = false;
        if (!attr_set) {
            OUT_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_bin_headers_lds),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kHeaderLds)));
            attr_set = true;
        }
        // about 4096 rows per workgroup: the zeroing and flush of the LDS bins
        // stay small next to the rows
        const int blocks = static_cast<int>(std::min<long long>((n_traces + 4095) / 4096, 1024));
        const int per_block = (n_traces + blocks - 1) / blocks;
        hipLaunchKernelGGL(k_bin_headers_lds, dim3(blocks), dim3(kHBlock), lds, st, source_x, source_y,
                           receiver_x, receiver_y, trace_index, n_traces, per_block, x_min, y_min, dx, dy, x_start,
                           y_start, tile_nx, tile_ny, grid_ny, trace_count, offset_sum, azimuth_sin_sum,
                           azimuth_cos_sum);
    } else {
        hipLaunchKernelGGL(k_bin_headers, dim3((n_traces + kBlock - 1) / kBlock), dim3(kBlock), 0, st, source_x,
                           source_y, receiver_x, receiver_y, trace_index, n_traces, x_min, y_min, dx, dy, x_start,
                           y_start, tile_nx, tile_ny, grid_ny, trace_count, offset_sum, azimuth_sin_sum,
                           azimuth_cos_sum);
    }
    OUT_TRY(hipGetLastError());
    return 0;
}

int pstm_hip_select_rows(const double* source_x, const double* source_y, const double* receiver_x,
                         const double* receiver_y, const int32_t* trace_index, int32_t n_traces, int32_t checks,
                         double offset_min, double offset_max, double ovt_x_min, double ovt_x_max, double ovt_y_min,
                         double ovt_y_max, uint8_t* keep, void* stream) {
    pstm_set_last_error("");
    if (n_traces < 0) return fail(1, "pstm_hip_select_rows: n_traces < 0");
    if (checks & ~0x3F) return fail(1, "pstm_hip_select_rows: unknown check bits");
    if (n_traces == 0) return 0;
    if (!source_x || !source_y || !receiver_x || !receiver_y || !keep)
        return fail(1, "pstm_hip_select_rows: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_select_rows, dim3((n_traces + kBlock - 1) / kBlock), dim3(kBlock), 0, st, source_x,
                       source_y, receiver_x, receiver_y, trace_index, n_traces, checks, offset_min, offset_max,
                       ovt_x_min, ovt_x_max, ovt_y_min, ovt_y_max, keep);
    OUT_TRY(hipGetLastError());
    return 0;
}

int pstm_hip_fold_normalize(const float* image, const int32_t* fold, long long n_pillars, int32_t nt, float* out,
                            void* stream) {
    pstm_set_last_error("");
    if (n_pillars < 0 || nt < 1) return fail(1, "pstm_hip_fold_normalize: n_pillars < 0 or nt < 1");
    if (n_pillars == 0) return 0;
    if (!image || !fold || !out) return fail(1, "pstm_hip_fold_normalize: null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const long long n = n_pillars * nt;
    const bool vec = nt % 4 == 0 && reinterpret_cast<uintptr_t>(image) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(out) % 16 == 0;
    if (vec) {
        hipLaunchKernelGGL(k_fold_normalize4, dim3(grid_for(n / 4)), dim3(kBlock), 0, st,
                           reinterpret_cast<const float4*>(image), fold, n / 4, nt / 4,
                           reinterpret_cast<float4*>(out));
    } else {
        hipLaunchKernelGGL(k_fold_normalize1, dim3(grid_for(n)), dim3(kBlock), 0, st, image, fold, n, nt, out);
    }
    OUT_TRY(hipGetLastError());
    return 0;
}

int pstm_hip_resample_uniform(const float* tv_image, const double* tv_times_ms, int32_t n_tv,
                              const double* out_times_ms, int32_t n_out, long long n_pillars, float* out,
                              void* stream) {
    pstm_set_last_error("");
    if (n_tv < 0 || n_out < 0 || n_pillars < 0) return fail(1, "pstm_hip_resample_uniform: negative size");
    if (n_out == 0 || n_pillars == 0) return 0;
    if (!out || !out_times_ms || (n_tv >= 2 && (!tv_image || !tv_times_ms)))
        return fail(1, "pstm_hip_resample_uniform: null pointer");
    const long long groups = (n_pillars + kRsPil - 1) / kRsPil;
    const int kblocks = (n_out + kBlock - 1) / kBlock;
    if (groups > 0x7FFFFFFFLL || kblocks > 65535) return fail(1, "pstm_hip_resample_uniform: volume too large");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t lds = n_tv <= kRsLds ? static_cast<size_t>(n_tv) * sizeof(double) : 0;
    hipLaunchKernelGGL(k_resample_uniform, dim3(static_cast<unsigned>(groups), kblocks), dim3(kBlock), lds, st,
                       tv_image, tv_times_ms, n_tv, out_times_ms, n_out, n_pillars, out);
    OUT_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
