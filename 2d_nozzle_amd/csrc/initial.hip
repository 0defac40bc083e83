// Point values of the quasi-1D initial state on the device (gfx950), behind
// nz_quasi1d_values in nozzle_hip.h: the step before the L2 projection
// (nz_project) that forms the march's initial state
// (src/timestepping/initialize.py:68-102).
//
// At each volume point (x, y): density, axial velocity and pressure
// interpolated linearly from the quasi-1D stations (numpy.interp's rules),
// the wall half-height y_wall(x) and its slope by the reference's central
// difference (h = 1e-6 L, one-sided at the inlet), the transverse velocity
// vx clip(y / y_wall, 0, 1) slope, and the conserved variables.
//
// The slope is a difference of wall heights 2e-6 apart, so one ulp in the
// wall moves it by about 1e-10.  The wall is therefore evaluated without FP
// contraction, in the host's operation order: for the Hermite (bell, smooth,
// moc) and conical walls every result is bit-identical to the host's.  The
// cubic and degree-9 Bezier walls need integer powers, which this file rounds
// correctly (double-double products); numpy's vectorised power is not always
// correctly rounded, so there the two can differ by an ulp.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "nozzle_hip.h"

namespace {

constexpr int BS = 256;
thread_local char g_ierr[512] = "";

int ifail(int code, const char* msg) {
    std::snprintf(g_ierr, sizeof g_ierr, "%s", msg);
    return code;
}

__device__ double clip01(double v) {
    const double a = v > 0.0 ? v : 0.0;  // numpy.clip: min(max(v, lo), hi)
    return a < 1.0 ? a : 1.0;
}

// s^k, rounded once from a double-double product (exact to ~2^-104)
__device__ double ipow(double s, int k) {
    if (k == 0) return 1.0;
    if (k == 1) return s;
    if (k == 2) return s * s;
    double hi = s, lo = 0.0;
    for (int i = 1; i < k; ++i) {
        const double p = hi * s;
        const double e = fma(hi, s, -p);  // exact error of hi * s
        const double l = fma(lo, s, e);
        hi = p + l;
        lo = l - (hi - p);
    }
    return hi + lo;
}

__device__ double hermite(double s, double y0, double y1, double m0, double m1) {
#pragma clang fp contract(off)
    const double s2 = s * s;
    const double s3 = s2 * s;
    return (2.0 * s3 - 3.0 * s2 + 1.0) * y0 + (s3 - 2.0 * s2 + s) * m0 +
           (-2.0 * s3 + 3.0 * s2) * y1 + (s3 - s2) * m1;
}

// wall half-height y_wall(x) (src/geometry/contour.py:264-311)
__device__ double wall(const nz_quasi1d_desc& d, double x) {
#pragma clang fp contract(off)
    const double xi = x / d.length;
    if (d.contour == NZ_CONTOUR_ANALYTIC) {
        const double t = d.pi * log1p(d.e2m1 * xi);
        const double half = cos(0.5 * t);
        return 0.01 * (4.0 - half) * (cos(t) + 5.0 - half);
    }
    const double x_th = d.throat_x;
    const double y_conv = hermite(clip01(xi / x_th), d.y_inlet, d.y_throat, 0.0, 0.0);
    const double span = 1.0 - x_th;
    const double s = clip01((xi - x_th) / span);
    double y_div;
    if (d.contour == NZ_CONTOUR_CONICAL) {
        y_div = d.y_throat + (d.y_exit - d.y_throat) * s;
    } else if (d.contour == NZ_CONTOUR_BEZIER9) {
        y_div = 0.0;
        for (int k = 0; k <= 9; ++k) {
            const double term = d.binom9[k] * ipow(s, k) * ipow(1.0 - s, 9 - k) * d.ctrl9[k];
            y_div = y_div + term;
        }
    } else if (d.contour == NZ_CONTOUR_BEZIER) {
        const double u = 1.0 - s;
        y_div = ipow(u, 3) * d.y_throat + 3.0 * ipow(u, 2) * s * d.bezier_c1 +
                3.0 * u * ipow(s, 2) * d.bezier_c2 + ipow(s, 3) * d.y_exit;
    } else {  // Hermite: bell, smooth, moc
        y_div = hermite(s, d.y_throat, d.y_exit, d.slope_throat, d.slope_exit);
    }
    return xi <= x_th ? y_conv : y_div;
}

// numpy.interp(x, xp, fp) for increasing xp: the left value below xp[0], the
// right value above xp[n-1], fp[j] on a station, else the secant through the
// bracketing stations evaluated from the left one
__device__ double interp(double x, const double* __restrict__ xp, const double* __restrict__ fp,
                         int n) {
#pragma clang fp contract(off)
    if (x < xp[0]) return fp[0];
    if (x > xp[n - 1]) return fp[n - 1];
    int lo = 0, hi = n - 1;  // xp[lo] <= x <= xp[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xp[mid] <= x) lo = mid;
        else hi = mid;
    }
    if (x == xp[n - 1]) return fp[n - 1];
    if (xp[lo] == x) return fp[lo];
    const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
    double r = slope * (x - xp[lo]) + fp[lo];
    if (r != r) r = slope * (x - xp[lo + 1]) + fp[lo + 1];
    return r;
}

__global__ __launch_bounds__(BS) void k_quasi1d_values(nz_quasi1d_desc d,
                                                       const double* __restrict__ xy, int64_t n,
                                                       double* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t i = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (i >= n) return;
    const double x = xy[2 * i], y = xy[2 * i + 1];
    const double rho = interp(x, d.x, d.density, d.n_points);
    const double vx = interp(x, d.x, d.velocity, d.n_points);
    const double p = interp(x, d.x, d.pressure, d.n_points);
    const double y_wall = wall(d, x);
    const double h = d.h;
    const double xm = x - h > 0.0 ? x - h : 0.0;
    const double slope = (wall(d, x + h) - wall(d, xm)) / (h + (x < h ? x : h));
    const double yw = y_wall > 1e-30 ? y_wall : 1e-30;
    const double vy = vx * clip01(y / yw) * slope;
    const double rhoE = p / (d.gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy);
    double* o = out + 4 * i;
    o[0] = rho;
    o[1] = rho * vx;
    o[2] = rho * vy;
    o[3] = rhoE;
}

}  // namespace

extern "C" {

const char* nz_initial_last_error(void) { return g_ierr; }

int nz_quasi1d_values(const nz_quasi1d_desc* d, const double* xy, int64_t n, double* values,
                      void* stream) {
    if (!d || !xy || !values) return ifail(NZ_E_ARG, "null argument");
    if (n < 0) return ifail(NZ_E_ARG, "negative point count");
    if (!d->x || !d->density || !d->velocity || !d->pressure || d->n_points < 2)
        return ifail(NZ_E_ARG, "the quasi-1D solution needs at least two stations");
    if (d->contour < NZ_CONTOUR_HERMITE || d->contour > NZ_CONTOUR_ANALYTIC)
        return ifail(NZ_E_ARG, "unknown contour id");
    if (!(d->length > 0.0)) return ifail(NZ_E_ARG, "length must be positive");
    if (n == 0) return NZ_OK;
    hipLaunchKernelGGL(k_quasi1d_values, dim3(static_cast<unsigned>((n + BS - 1) / BS)), dim3(BS),
                       0, static_cast<hipStream_t>(stream), *d, xy, n, values);
    const hipError_t st = hipGetLastError();
    if (st != hipSuccess) {
        std::snprintf(g_ierr, sizeof g_ierr, "nz_quasi1d_values: %s", hipGetErrorString(st));
        return NZ_E_LAUNCH;
    }
    return NZ_OK;
}

}  // extern "C"
