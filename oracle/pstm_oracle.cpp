// pstm_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the reference PSTM tile-migration kernel, used as the
// correctness checker of the HIP backend and as the CPU baseline timed beside
// it.  It is never linked into, loaded by or called from the product library
// (pm_amd/); only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it.
//
// What it follows (reference paths relative to the Nordavind82/PM checkout):
//   * _precompute_time_terms        pstm/kernels/numba_cpu_optimized.py:358-401
//   * _migrate_tile_kernel_optimized pstm/kernels/numba_cpu_optimized.py:93-233
//     (loop nest: pillars in parallel, traces in index order, output samples;
//      fold = per-pillar count of traces with midpoint distance <= max
//      aperture, added to every sample of the pillar, :154,171-175,231)
//   * 3-D velocity v[ix,iy,it] indexing   pstm/metal/shaders/pstm_migration.metal:187-197
//     (the registered Numba kernel has no 3-D path; this is the build's own
//      extension with the aperture taken from the local velocity)
//   * curved-ray traveltime   pstm/algorithm/curved_ray.py:19-52,239-262,287-316;
//     weight velocity v0+k z   pstm/metal/shaders/pstm_curved_ray.metal:209,268-270
//   * VTI traveltime          pstm/algorithm/anisotropy_vti.py:148-220
//   * trace interpolation other than linear: interpolate_sample and its kernels
//     (pstm/kernels/interpolation.py:58-318), as the reference's NumbaKernel
//     applies it (pstm/kernels/numba_cpu.py:337-341, 632)
//   * triangle anti-aliasing on double-integrated traces: NOT in the reference
//     (its Metal shaders use a scalar weight, pstm_migration.metal:95-132).
//     Method of Lumley, Claerbout & Bevc (1994): a triangle of half width L
//     samples, L from the operator time dip, evaluated as second differences
//     (P[n+L] - 2P[n] + P[n-L]) / L^2 of the double running sum P; fractional
//     L blends the two neighbouring integer widths (triangle_at).
//
// Parity pins (tests/test_oracle.py): golden vectors written from the
// reference's own functions (tests/golden/make_golden.py), closed forms,
// brute-force NumPy loop nests and explicit triangle convolutions.
//
// Build: see oracle/Makefile (strict IEEE build and a -ffast-math build).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/pstm_hip.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Double running sum P of one trace A[0..ns-1] with the sample convention
// P[n] = sum_{m<n} S[m], S[m] = sum_{q<=m} A[q]; P[n] = 0 for n <= 0 and
// P[n] = P[ns] + (n - ns) * S[ns-1] for n > ns (A = 0 outside the trace).
// Then P[n+1] - 2 P[n] + P[n-1] = A[n] exactly in exact arithmetic.
struct Integrated {
    std::vector<double> p;  // P[0..ns]
    double s_end = 0.0;     // S[ns-1]
};

void integrate_trace(const float* a, int ns, Integrated& out) {
    out.p.assign(ns + 1, 0.0);
    double s = 0.0, p = 0.0;
    for (int n = 0; n < ns; ++n) {
        s += (double)a[n];
        p += s;
        out.p[n + 1] = p;
    }
    out.s_end = s;
}

inline double p_at(const Integrated& I, long n, int ns) {
    if (n <= 0) return 0.0;
    if (n <= ns) return I.p[n];
    return I.p[ns] + (double)(n - ns) * I.s_end;
}

// Triangle filter of integer half width L at integer sample n:
// (P[n+L] - 2 P[n] + P[n-L]) / L^2 = (1/L^2) sum_{|m|<L} (L-|m|) A[n+m].
inline double tri_int(const Integrated& I, long n, long L, int ns) {
    return (p_at(I, n + L, ns) - 2.0 * p_at(I, n, ns) + p_at(I, n - L, ns)) / (double)(L * L);
}

// Triangle of fractional half width L >= 1 at fractional position u: linear
// blend of the integer half widths floor(L) and floor(L)+1 (weight L -
// floor(L)), each linearly interpolated between the samples around u.  Unit
// DC gain for every L and u, continuous in L and u; L = 1 is plain linear
// interpolation of the trace.
inline double triangle_at(const Integrated& I, double u, double L, int ns) {
    const long i0 = (long)std::floor(u);
    const double f = u - (double)i0;
    const long Lf = (long)std::floor(L);
    const double lam = L - (double)Lf;
    const double t0 = (1.0 - lam) * tri_int(I, i0, Lf, ns) + lam * tri_int(I, i0, Lf + 1, ns);
    const double t1 = (1.0 - lam) * tri_int(I, i0 + 1, Lf, ns) + lam * tri_int(I, i0 + 1, Lf + 1, ns);
    return t0 + f * (t1 - t0);
}

// curved_ray.py:239-262
inline double time_to_depth_curved(double t_s, double v0, double k) {
    if (std::fabs(k) < 1e-10) return v0 * t_s;
    return (v0 / k) * (std::exp(k * t_s) - 1.0);
}

// curved_ray.py:19-52
inline double curved_ray_traveltime(double x, double z, double v0, double k) {
    if (std::fabs(k) < 1e-10) {
        double r = std::sqrt(x * x + z * z);
        return r > 0 ? r / v0 : 0.0;
    }
    double v_z = v0 + k * z;
    double kx = k * x;
    double term1 = v_z + std::sqrt(v_z * v_z + kx * kx);
    double term2 = v0 + std::sqrt(v0 * v0 + kx * kx);
    if (term2 <= 0) return 0.0;
    return (1.0 / k) * std::log(term1 / term2);
}

// anisotropy_vti.py:181-220
inline double vti_one_way_time(double t0_one_way, double h, double v_nmo, double eta) {
    if (v_nmo <= 0 || t0_one_way <= 0) return t0_one_way;
    double h2 = h * h;
    double v2 = v_nmo * v_nmo;
    double t02 = t0_one_way * t0_one_way;
    double t2_hyper = t02 + h2 / v2;
    double denom = v2 * (t02 * v2 + (1.0 + 2.0 * eta) * h2);
    if (std::fabs(denom) < 1e-20) return std::sqrt(std::fmax(t2_hyper, 0.0));
    double correction = (2.0 * eta * h2 * h2) / denom;
    double t2 = t2_hyper - correction;
    return std::sqrt(std::fmax(t2, 0.0));
}

// Trace interpolation of the reference's NumbaKernel (interpolate_sample,
// pstm/kernels/interpolation.py:58-318) in f64 on an f32 trace, codes
// PSTM_INTERP_* (the reference's get_method_code numbers with linear and
// nearest swapped).
inline double interp_linear(const float* tr, int ns, double t) {  // :68-92
    if (t < 0.0 || t >= ns - 1) return 0.0;
    const int i0 = (int)t;
    const int i1 = i0 + 1;
    if (i1 >= ns) return tr[i0];
    const double frac = t - i0;
    return (double)tr[i0] * (1.0 - frac) + (double)tr[i1] * frac;
}

inline double sinc_value(double x) {  // :138-143
    if (std::fabs(x) < 1e-10) return 1.0;
    return std::sin(M_PI * x) / (M_PI * x);
}

double interp_sinc(const float* tr, int ns, double t, int hw) {  // :146-191
    if (t < hw || t >= ns - hw) return interp_linear(tr, ns, t);
    const int ic = (int)(t + 0.5);
    const double frac = t - ic;
    double res = 0.0, norm = 0.0;
    for (int k = -hw + 1; k <= hw; ++k) {
        const int i = ic + k;
        if (i < 0 || i >= ns) continue;
        const double x = frac - k;
        const double win = std::fabs(x) >= hw ? 0.0 : 0.5 * (1.0 + std::cos(M_PI * x / hw));
        const double w = sinc_value(x) * win;
        res += (double)tr[i] * w;
        norm += w;
    }
    return norm > 1e-10 ? res / norm : 0.0;
}

double interp_lanczos(const float* tr, int ns, double t, int a) {  // :212-256
    if (t < a || t >= ns - a) return interp_linear(tr, ns, t);
    const int ic = (int)t;
    const double frac = t - ic;
    double res = 0.0, norm = 0.0;
    for (int k = -a + 1; k <= a; ++k) {
        const int i = ic + k;
        if (i < 0 || i >= ns) continue;
        const double x = frac - k;
        double w;
        if (std::fabs(x) < 1e-10) w = 1.0;
        else if (std::fabs(x) >= a) w = 0.0;
        else w = sinc_value(x) * sinc_value(x / a);
        res += (double)tr[i] * w;
        norm += w;
    }
    return norm > 1e-10 ? res / norm : 0.0;
}

double interp_sample(const float* tr, int ns, double t, int method) {  // :276-318
    switch (method) {
        case PSTM_INTERP_NEAREST: {  // :58-65
            const int i = (int)(t + 0.5);
            return (i < 0 || i >= ns) ? 0.0 : (double)tr[i];
        }
        case PSTM_INTERP_CUBIC: {  // Catmull-Rom, :95-135
            if (t < 1.0 || t >= ns - 2) return interp_linear(tr, ns, t);
            const int i = (int)t;
            const double u = t - i;
            const double p0 = tr[i - 1], p1 = tr[i], p2 = tr[i + 1], p3 = tr[i + 2];
            const double u2 = u * u, u3 = u2 * u;
            const double a0 = -0.5 * p0 + 1.5 * p1 - 1.5 * p2 + 0.5 * p3;
            const double a1 = p0 - 2.5 * p1 + 2.0 * p2 - 0.5 * p3;
            const double a2 = -0.5 * p0 + 0.5 * p2;
            return a0 * u3 + a1 * u2 + a2 * u + p1;
        }
        case PSTM_INTERP_SINC4: return interp_sinc(tr, ns, t, 2);
        case PSTM_INTERP_SINC8: return interp_sinc(tr, ns, t, 4);
        case PSTM_INTERP_SINC16: return interp_sinc(tr, ns, t, 8);
        case PSTM_INTERP_LANCZOS3: return interp_lanczos(tr, ns, t, 3);
        case PSTM_INTERP_LANCZOS5: return interp_lanczos(tr, ns, t, 5);
        default: return interp_linear(tr, ns, t);
    }
}

}  // namespace

extern "C" {

// One interpolated sample, for the golden-vector checks.
double oracle_interp_sample(const float* trace, int ns, double t, int method) {
    return interp_sample(trace, ns, t, method);
}

int oracle_abi_version(void) { return PSTM_HIP_ABI_VERSION; }
size_t oracle_args_size(void) { return sizeof(PstmMigrateArgs); }

// Host-pointer variant of pstm_hip_migrate_tile.  `image64` (nx,ny,nt) f64 and
// a->fold (nx,ny) are accumulated; a->image is ignored.  *n_effective (if
// non-NULL) receives E, the contributions that passed the time-dependent
// aperture and the sample-bounds test.
int oracle_migrate_tile(const PstmMigrateArgs* a, double* image64, int num_threads) {
    const int nx = a->nx, ny = a->ny, nt = a->nt, ns = a->n_samples;
    const int ntr = a->n_traces;
    const long n_pillars = (long)nx * ny;
    const bool vel3d = a->vel_mode == PSTM_VEL_3D;
    const double dt_in_s = a->dt_in_ms / 1000.0;
    const double t_start_in_s = a->t_start_in_ms / 1000.0;
    const double max_ap = a->max_aperture_m, min_ap = a->min_aperture_m;
    const double tf = a->taper_fraction;
    const double tan_dip = std::tan(a->max_dip_deg * M_PI / 180.0);
    const int tt = a->traveltime;
    const bool aa = a->aa_mode == PSTM_AA_TRIANGLE;
    const double lmax = a->aa_max_half_width > 0 ? (double)a->aa_max_half_width : 1.0;
    if (aa && tt != PSTM_TT_STRAIGHT) return 2;

    // _precompute_time_terms (numba_cpu_optimized.py:375-401); per pillar when
    // the velocity is 3-D.
    const long n_vel = vel3d ? n_pillars : 1;
    std::vector<double> t0s(nt), q(nt), s((size_t)n_vel * nt), ap((size_t)n_vel * nt);
    for (int k = 0; k < nt; ++k) {
        double t_s = a->t_axis_ms[k] / 1000.0;
        t0s[k] = t_s;
        q[k] = (t_s * 0.5) * (t_s * 0.5);
    }
    for (long p = 0; p < n_vel; ++p)
        for (int k = 0; k < nt; ++k) {
            double v = a->vrms[(size_t)p * nt + k];
            double t_s = t0s[k];
            s[(size_t)p * nt + k] = 1.0 / (v * v);
            double apv = v * t_s * 0.5 * tan_dip;
            if (apv < min_ap) apv = min_ap;
            else if (apv > max_ap) apv = max_ap;
            ap[(size_t)p * nt + k] = apv;
        }
    // curved ray: depth and velocity at depth per output time
    std::vector<double> zc(nt), vz(nt);
    if (tt == PSTM_TT_CURVED)
        for (int k = 0; k < nt; ++k) {
            zc[k] = time_to_depth_curved(t0s[k] / 2.0, a->curved_v0, a->curved_k);
            vz[k] = a->curved_v0 + a->curved_k * zc[k];
        }

    // integrated traces for anti-aliasing
    std::vector<Integrated> integ;
    if (aa) {
        integ.resize(ntr);
#pragma omp parallel for schedule(static) num_threads(num_threads > 0 ? num_threads : 1)
        for (int jj = 0; jj < ntr; ++jj) {
            int j = a->trace_index ? a->trace_index[jj] : jj;
            integrate_trace(a->amplitudes + (size_t)j * ns, ns, integ[jj]);
        }
    }

    unsigned long long e_total = 0;
    const int nthr = num_threads > 0 ? num_threads : 1;
#pragma omp parallel for schedule(static) num_threads(nthr) reduction(+ : e_total)
    for (long idx = 0; idx < n_pillars; ++idx) {
        const int ix = (int)(idx / ny), iy = (int)(idx % ny);
        const double ox = a->x_axis[ix], oy = a->y_axis[iy];
        const size_t vb = vel3d ? (size_t)idx * nt : 0;
        double* img = image64 + (size_t)idx * nt;
        int local_fold = 0;
        unsigned long long e_local = 0;
        for (int jj = 0; jj < ntr; ++jj) {
            const int j = a->trace_index ? a->trace_index[jj] : jj;
            const double sx = a->source_x[j], sy = a->source_y[j];
            const double rx = a->receiver_x[j], ry = a->receiver_y[j];
            const double mx = a->midpoint_x[j], my = a->midpoint_y[j];
            const double w_trace = a->trace_weights ? a->trace_weights[j] : 1.0;
            const double dxm = ox - mx, dym = oy - my;
            const double dist_mid = std::sqrt(dxm * dxm + dym * dym);
            if (dist_mid > max_ap) continue;
            local_fold += 1;
            const double dxs = ox - sx, dys = oy - sy, dxr = ox - rx, dyr = oy - ry;
            const double ds2 = dxs * dxs + dys * dys;
            const double dr2 = dxr * dxr + dyr * dyr;
            const float* trace = a->amplitudes + (size_t)j * ns;
            for (int k = 0; k < nt; ++k) {
                const double aperture = ap[vb + k];
                if (dist_mid > aperture) continue;
                const double sk = s[vb + k];
                double t_total_s, t_src = 0, t_rec = 0, v_weight;
                if (tt == PSTM_TT_STRAIGHT) {
                    t_src = std::sqrt(q[k] + ds2 * sk);
                    t_rec = std::sqrt(q[k] + dr2 * sk);
                    t_total_s = t_src + t_rec;
                    v_weight = 1.0 / std::sqrt(sk);
                } else if (tt == PSTM_TT_CURVED) {
                    t_total_s = curved_ray_traveltime(std::sqrt(ds2), zc[k], a->curved_v0, a->curved_k) +
                                curved_ray_traveltime(std::sqrt(dr2), zc[k], a->curved_v0, a->curved_k);
                    v_weight = vz[k];
                } else {
                    double eta;
                    if (a->eta_mode == PSTM_ETA_1D) eta = a->eta[k];
                    else if (a->eta_mode == PSTM_ETA_3D) eta = a->eta[(size_t)idx * nt + k];
                    else eta = a->eta_constant;
                    const double v = 1.0 / std::sqrt(sk);
                    const double tau = t0s[k] / 2.0;
                    t_total_s = vti_one_way_time(tau, std::sqrt(ds2), v, eta) +
                                vti_one_way_time(tau, std::sqrt(dr2), v, eta);
                    v_weight = v;
                }
                const double t_sample = (t_total_s - t_start_in_s) / dt_in_s;
                if (t_sample < 0.0 || t_sample >= ns - 1) continue;
                double amp;
                if (!aa && a->interpolation != PSTM_INTERP_LINEAR) {
                    amp = interp_sample(trace, ns, t_sample, a->interpolation);
                } else if (!aa) {
                    const int i0 = (int)t_sample;
                    const double frac = t_sample - i0;
                    amp = (double)trace[i0] * (1.0 - frac) + (double)trace[i0 + 1] * frac;
                } else {
                    // operator time dip (s per m) along x and y
                    double gx = 0.0, gy = 0.0;
                    if (t_src > 0) { gx += dxs * sk / t_src; gy += dys * sk / t_src; }
                    if (t_rec > 0) { gx += dxr * sk / t_rec; gy += dyr * sk / t_rec; }
                    double L = std::fmax(std::fabs(gx) * a->grid_dx, std::fabs(gy) * a->grid_dy) / dt_in_s;
                    if (!(L > 1.0)) L = 1.0;
                    if (L > lmax) L = lmax;
                    amp = triangle_at(integ[jj], t_sample, L, ns);
                }
                double weight = w_trace;
                if (dist_mid > aperture * (1.0 - tf)) {
                    const double taper_start = aperture * (1.0 - tf);
                    const double x = (dist_mid - taper_start) / (aperture - taper_start);
                    weight *= 0.5 * (1.0 + std::cos(M_PI * x));
                }
                if (a->apply_spreading && t_total_s > 0.001) weight *= 1.0 / (v_weight * t_total_s);
                if (a->apply_obliquity && t_total_s > 0.001) weight *= t0s[k] / t_total_s;
                img[k] += amp * weight;
                e_local += 1;
            }
        }
        a->fold[idx] += local_fold;
        e_total += e_local;
    }
    if (a->n_effective) *a->n_effective += e_total;
    return 0;
}

// Fold only: the per-pillar count of traces with midpoint distance <= max
// aperture, added to a->fold (numba_cpu_optimized.py:154,167-175,231), the
// same f64 test as oracle_migrate_tile's; for checking full-size jobs whose
// whole image is too costly to recompute on the CPU.
int oracle_fold_tile(const PstmMigrateArgs* a, int num_threads) {
    const int nx = a->nx, ny = a->ny;
    const int ntr = a->n_traces;
    const double max_ap = a->max_aperture_m;
    const int nthr = num_threads > 0 ? num_threads : 1;
#pragma omp parallel for schedule(static) num_threads(nthr)
    for (long idx = 0; idx < (long)nx * ny; ++idx) {
        const double ox = a->x_axis[idx / ny], oy = a->y_axis[idx % ny];
        int lf = 0;
        for (int jj = 0; jj < ntr; ++jj) {
            const int j = a->trace_index ? a->trace_index[jj] : jj;
            const double dxm = ox - a->midpoint_x[j], dym = oy - a->midpoint_y[j];
            if (!(std::sqrt(dxm * dxm + dym * dym) > max_ap)) ++lf;
        }
        a->fold[idx] += lf;
    }
    return 0;
}

// Double running sums of `n_traces` traces (rows of amp): p (n_traces, ns+1)
// = P[0..ns] and s_end (n_traces,) = S[ns-1], the convention of
// integrate_trace above.
void oracle_integrate_traces(const float* amp, int n_traces, int ns, double* p, double* s_end) {
    for (int j = 0; j < n_traces; ++j) {
        Integrated I;
        integrate_trace(amp + (size_t)j * ns, ns, I);
        std::memcpy(p + (size_t)j * (ns + 1), I.p.data(), sizeof(double) * (ns + 1));
        s_end[j] = I.s_end;
    }
}

// Triangle filter of half width L at position u on one trace (the
// anti-aliased sample of a contribution), for the filter checks.
double oracle_triangle_at(const float* amp, int ns, double u, double L) {
    Integrated I;
    integrate_trace(amp, ns, I);
    return triangle_at(I, u, L, ns);
}

// Scalar traveltimes, for the closed-form and golden checks.
double oracle_curved_dsr(double t0_s, double h_s, double h_r, double v0, double k) {
    double z = time_to_depth_curved(t0_s / 2, v0, k);
    return curved_ray_traveltime(h_s, z, v0, k) + curved_ray_traveltime(h_r, z, v0, k);
}
double oracle_vti_dsr(double t0_s, double h_s, double h_r, double v, double eta) {
    double t0_half = t0_s / 2.0;
    return vti_one_way_time(t0_half, h_s, v, eta) + vti_one_way_time(t0_half, h_r, v, eta);
}
double oracle_vti_leg(double tau, double h, double v, double eta) {
    return vti_one_way_time(tau, h, v, eta);
}

}  // extern "C"
