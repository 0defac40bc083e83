// Observables of a state on the device (gfx950), behind
// nz_performance_integrals and nz_objective in nozzle_hip.h.
//
//   * the volume and exit integrals behind the performance fields that the
//     boundary integrals do not give (src/post/postprocess.py:174-255): the
//     entropy error's integral of (p / (rho^gamma s_t) - 1)^2 over the domain
//     and the exit integral of p / rho;
//   * the objectives of the adjoint (src/sensitivity.py:96-130): thrust,
//     thrust coefficient, mass flow, exit Mach and exit pressure, with their
//     partial derivatives with respect to the state (E, N, 4) and to the
//     edge geometry (n_x, n_y, |dX/dsigma|) per boundary edge point.
//
// Every sum runs in a fixed order (per-point values, then one block that
// walks them in index order), and the state cotangent is gathered per
// element from its own boundary faces, so results are reproducible bit for
// bit.  These run once per solve (or per gradient): generic loops over N,
// nqv and nqf rather than per-configuration instantiations.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "nozzle_hip.h"

namespace {

constexpr int BS = 256;
constexpr double kFloor = 1e-10;
thread_local char g_oerr[512] = "";

int ofail(int code, const char* msg) {
    std::snprintf(g_oerr, sizeof g_oerr, "%s", msg);
    return code;
}

int ocheck(const char* what) {
    const hipError_t st = hipGetLastError();
    if (st != hipSuccess) {
        std::snprintf(g_oerr, sizeof g_oerr, "%s: %s", what, hipGetErrorString(st));
        return NZ_E_LAUNCH;
    }
    return NZ_OK;
}

inline unsigned nblk(int64_t n) { return static_cast<unsigned>((n + BS - 1) / BS); }

// fixed-order sum of x[0..n) (stride `stride`, offset `off`) by one block
__device__ double block_fixed_sum(const double* __restrict__ x, int64_t n, int stride, int off,
                                  double* red) {
    const int t = threadIdx.x;
    double a = 0.0;
    for (int64_t i = t; i < n; i += BS) a += x[i * stride + off];
    red[t] = a;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------ performance
// One thread per element: sum_q w det (p / (rho^gamma s_t) - 1)^2 at the
// volume points (entropy_error, postprocess.py:174-190).
__global__ void k_entropy_elements(nz_problem P, const double* __restrict__ U, double s_t,
                                   double* __restrict__ part) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (e >= P.n_elem) return;
    const double* T = P.tables;
    const int N = P.nbf, nqv = P.nqv;
    double acc = 0.0;
    for (int q = 0; q < nqv; ++q) {
        double S[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < N; ++i) {
            const double ph = T[P.off.phi_vol + i * nqv + q];
            for (int c = 0; c < 4; ++c) S[c] += ph * U[(e * N + i) * 4 + c];
        }
        const double r = fmax(S[0], kFloor);
        const double vx = S[1] / r, vy = S[2] / r;
        const double p = fmax((P.gamma - 1.0) * (S[3] - 0.5 * r * (vx * vx + vy * vy)), kFloor);
        const double d = p / pow(r, P.gamma) / s_t - 1.0;
        const double wd = P.affine ? T[P.off.w_vol + q] * P.geo[e * 5 + 4]
                                   : P.geo[(e * nqv + q) * 5 + 4];
        acc += wd * d * d;
    }
    part[e] = acc;
}

// One thread per outflow point: p / rho ds (the exit temperature times R).
__global__ void k_exit_temperature_points(nz_problem P, const double* __restrict__ U,
                                          double* __restrict__ val) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int nqf = P.nqf, N = P.nbf;
    const int64_t nb = P.n_edge - P.n_int;
    if (it >= nb * nqf) return;
    const int64_t b = it / nqf;
    const int q = static_cast<int>(it - b * nqf);
    const int ft = P.bedge[2 * b + 1];
    const int f = ft >> 3, tag = ft & 7;
    if (tag != NZ_TAG_OUTFLOW) {
        val[it] = 0.0;
        return;
    }
    const int64_t e = P.bedge[2 * b];
    const double* T = P.tables;
    double S[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < N; ++i) {
        const double w = T[P.off.phi_face + (f * N + i) * nqf + q];
        for (int c = 0; c < 4; ++c) S[c] += w * U[(e * N + i) * 4 + c];
    }
    const int64_t k = P.n_int + b;
    const double jac = P.straight ? P.edge_geo[k * 3 + 2] : P.edge_geo[(k * nqf + q) * 3 + 2];
    const double ds = jac * T[P.off.w_face + q];
    const double r = fmax(S[0], kFloor), vx = S[1] / r, vy = S[2] / r;
    const double p = fmax((P.gamma - 1.0) * (S[3] - 0.5 * r * (vx * vx + vy * vy)), kFloor);
    val[it] = p / r * ds;
}

__global__ void k_two_sums(const double* __restrict__ a, int64_t na, const double* __restrict__ b,
                           int64_t nbv, double* __restrict__ out) {
    __shared__ double red[BS];
    const double sa = block_fixed_sum(a, na, 1, 0, red);
    const double sb = block_fixed_sum(b, nbv, 1, 0, red);
    if (threadIdx.x == 0) {
        out[0] = sa;
        out[1] = sb;
    }
}

// ------------------------------------------------------------- objectives
struct PointState {
    double rho, drho, m1, m2, ke, p, dp_on, dp[4];
};

__device__ PointState point_state(const double S[4], double g) {
    PointState s;
    const double rho_raw = S[0];
    s.rho = fmax(rho_raw, kFloor);
    s.drho = rho_raw > kFloor ? 1.0 : 0.0;
    s.m1 = S[1];
    s.m2 = S[2];
    s.ke = 0.5 * (s.m1 * s.m1 + s.m2 * s.m2) / s.rho;
    const double p_raw = (g - 1.0) * (S[3] - s.ke);
    s.p = fmax(p_raw, kFloor);
    s.dp_on = p_raw > kFloor ? 1.0 : 0.0;
    s.dp[0] = (g - 1.0) * s.ke / s.rho * s.drho * s.dp_on;
    s.dp[1] = -(g - 1.0) * s.m1 / s.rho * s.dp_on;
    s.dp[2] = -(g - 1.0) * s.m2 / s.rho * s.dp_on;
    s.dp[3] = (g - 1.0) * s.dp_on;
    return s;
}

// One thread per boundary point.  Integral objectives write val = J's
// integrand times ds (with the objective's factor), gS = its state
// derivative times ds and the edge-geometry derivatives into dG; the
// averaged objectives (exit Mach, exit pressure) write the numerator's
// integrand and ds, and their derivatives are finished once the averages
// are known.
__global__ void k_objective_points(nz_problem P, const double* __restrict__ U, int obj,
                                   double scale, double* __restrict__ val, double* __restrict__ gS,
                                   double* __restrict__ dG) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int nqf = P.nqf, N = P.nbf;
    const int64_t nb = P.n_edge - P.n_int;
    if (it >= nb * nqf) return;
    const int64_t b = it / nqf;
    const int q = static_cast<int>(it - b * nqf);
    const int ft = P.bedge[2 * b + 1];
    const int f = ft >> 3, tag = ft & 7;
    const int64_t k = P.n_int + b;
    double* g = dG + (k * nqf + q) * 3;
    double* gs = gS + it * 4;
    const bool thrust = obj == NZ_OBJ_THRUST || obj == NZ_OBJ_THRUST_COEFFICIENT;
    const bool mass = obj == NZ_OBJ_MASS_FLOW;
    const bool on = thrust ? (tag == NZ_TAG_OUTFLOW || tag == NZ_TAG_INFLOW)
                           : (mass ? tag == NZ_TAG_INFLOW : tag == NZ_TAG_OUTFLOW);
    val[2 * it] = 0.0;
    val[2 * it + 1] = 0.0;
    for (int c = 0; c < 4; ++c) gs[c] = 0.0;
    g[0] = g[1] = g[2] = 0.0;
    if (!on) return;
    const int64_t e = P.bedge[2 * b];
    const double* T = P.tables;
    double S[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < N; ++i) {
        const double w = T[P.off.phi_face + (f * N + i) * nqf + q];
        for (int c = 0; c < 4; ++c) S[c] += w * U[(e * N + i) * 4 + c];
    }
    const double* eg = P.straight ? P.edge_geo + k * 3 : P.edge_geo + (k * nqf + q) * 3;
    const double nx = eg[0], ny = eg[1], wq = T[P.off.w_face + q];
    const double ds = eg[2] * wq;
    const PointState s = point_state(S, P.gamma);
    const double rho = s.rho, m1 = s.m1, m2 = s.m2;
    if (thrust || mass) {
        const double fac = (thrust ? 2.0 : -2.0) * scale;
        const double vn = m1 * nx + m2 * ny;  // rho v_n
        double I, dI[4];
        if (thrust) {
            I = m1 * vn / rho + s.p * nx;
            dI[0] = -m1 * vn / (rho * rho) * s.drho + s.dp[0] * nx;
            dI[1] = (vn + m1 * nx) / rho + s.dp[1] * nx;
            dI[2] = m1 * ny / rho + s.dp[2] * nx;
            dI[3] = s.dp[3] * nx;
            g[0] = fac * (m1 * m1 / rho + s.p) * ds;
            g[1] = fac * (m1 * m2 / rho) * ds;
        } else {
            I = vn;
            dI[0] = 0.0;
            dI[1] = nx;
            dI[2] = ny;
            dI[3] = 0.0;
            g[0] = fac * m1 * ds;
            g[1] = fac * m2 * ds;
        }
        g[2] = fac * I * wq;
        val[2 * it] = fac * I * ds;
        for (int c = 0; c < 4; ++c) gs[c] = fac * dI[c] * ds;
        return;
    }
    // averaged over the outflow: exit Mach, or exit pressure over p_t (scale)
    double Q, dQ[4];
    if (obj == NZ_OBJ_EXIT_MACH) {
        const double q2 = m1 * m1 + m2 * m2;
        // M^2 = (m1^2 + m2^2) / (rho^2) * rho / (gamma p)
        const double vx = m1 / rho, vy = m2 / rho;
        Q = sqrt((vx * vx + vy * vy) * rho / (P.gamma * s.p));
        if (q2 > 0.0) {
            const double h = 0.5 * Q;
            dQ[0] = h * (-s.drho / rho - s.dp[0] / s.p);
            dQ[1] = h * (2.0 * m1 / q2 - s.dp[1] / s.p);
            dQ[2] = h * (2.0 * m2 / q2 - s.dp[2] / s.p);
            dQ[3] = h * (-s.dp[3] / s.p);
        } else {
            dQ[0] = dQ[1] = dQ[2] = dQ[3] = 0.0;
        }
    } else {
        Q = s.p * scale;
        for (int c = 0; c < 4; ++c) dQ[c] = s.dp[c] * scale;
    }
    val[2 * it] = Q * ds;
    val[2 * it + 1] = ds;
    for (int c = 0; c < 4; ++c) gs[c] = dQ[c] * ds;  // divided by the area below
    g[2] = Q;                                          // (Q - J) w / area below
}

__global__ void k_objective_reduce(const double* __restrict__ val, int64_t n, int averaged,
                                   double* __restrict__ out) {
    __shared__ double red[BS];
    const double a = block_fixed_sum(val, n, 2, 0, red);
    const double b = block_fixed_sum(val, n, 2, 1, red);
    if (threadIdx.x == 0) {
        out[0] = averaged ? a / b : a;
        out[1] = b;
    }
}

// averaged objectives: gS /= area; d/d jac = (Q - J) w / area
__global__ void k_objective_average(nz_problem P, const double* __restrict__ red,
                                    double* __restrict__ gS, double* __restrict__ dG) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int nqf = P.nqf;
    const int64_t nb = P.n_edge - P.n_int;
    if (it >= nb * nqf) return;
    const int64_t b = it / nqf;
    if ((P.bedge[2 * b + 1] & 7) != NZ_TAG_OUTFLOW) return;  // not in the average: zeros
    const double J = red[0], area = red[1];
    const int q = static_cast<int>(it - b * nqf);
    double* g = dG + ((P.n_int + b) * nqf + q) * 3;
    double* gs = gS + it * 4;
    const double wq = P.tables[P.off.w_face + q];
    for (int c = 0; c < 4; ++c) gs[c] = gs[c] / area;
    g[2] = (g[2] - J) * wq / area;  // g[2] held Q
}

// One thread per element: dJ/dU = sum over its boundary faces of
// phi_face(side 0)^T gS, faces in order (no atomics).
__global__ void k_objective_elements(nz_problem P, const double* __restrict__ gS,
                                     double* __restrict__ dU) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (e >= P.n_elem) return;
    const int N = P.nbf, nqf = P.nqf, nf = P.nface;
    const double* T = P.tables;
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 4; ++c) dU[(e * N + i) * 4 + c] = 0.0;
    for (int f = 0; f < nf; ++f) {
        if (P.fe_nbr[e * nf + f] >= 0) continue;
        const int64_t b = static_cast<int64_t>(P.fe_info[e * nf + f] >> 3) - P.n_int;
        const double* gs = gS + b * nqf * 4;
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 4; ++c) {
                double a = dU[(e * N + i) * 4 + c];
                for (int q = 0; q < nqf; ++q)
                    a += T[P.off.phi_face + (f * N + i) * nqf + q] * gs[q * 4 + c];
                dU[(e * N + i) * 4 + c] = a;
            }
    }
}

__global__ void k_zero_interior_edges(nz_problem P, double* __restrict__ dG) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (it < P.n_int * P.nqf * 3) dG[it] = 0.0;
}

int check_problem(const nz_problem* p) {
    if (!p || p->n_elem <= 0 || p->nbf <= 0 || p->nqf <= 0 || p->nqv <= 0)
        return ofail(NZ_E_ARG, "invalid problem");
    if (p->n_int < 0 || p->n_int > p->n_edge) return ofail(NZ_E_ARG, "n_int out of range");
    if (!p->tables || !p->geo || !p->edge_geo || !p->fe_nbr || !p->fe_info ||
        (p->n_edge > p->n_int && !p->bedge))
        return ofail(NZ_E_ARG, "problem arrays missing");
    return NZ_OK;
}

}  // namespace

extern "C" {

const char* nz_observables_last_error(void) { return g_oerr; }

int64_t nz_observables_work_doubles(const nz_problem* p) {
    if (check_problem(p)) return -1;
    const int64_t nbq = (p->n_edge - p->n_int) * p->nqf;
    const int64_t a = p->n_elem > 2 * nbq ? p->n_elem : 2 * nbq;
    return a + 4 * nbq + 8;
}

int nz_performance_integrals(const nz_problem* p, const double* U, double s_t, double* work,
                             double* out, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!U || !work || !out) return ofail(NZ_E_ARG, "null array");
    if (!(s_t > 0.0)) return ofail(NZ_E_ARG, "s_t must be positive");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nbq = (p->n_edge - p->n_int) * p->nqf;
    double* part = work;             // (E,)
    double* pts = work + p->n_elem;  // (nbq,)
    hipLaunchKernelGGL(k_entropy_elements, dim3(nblk(p->n_elem)), dim3(BS), 0, st, *p, U, s_t,
                       part);
    if (nbq > 0)
        hipLaunchKernelGGL(k_exit_temperature_points, dim3(nblk(nbq)), dim3(BS), 0, st, *p, U,
                           pts);
    hipLaunchKernelGGL(k_two_sums, dim3(1), dim3(BS), 0, st, part, p->n_elem, pts, nbq, out);
    return ocheck("nz_performance_integrals");
}

int nz_objective(const nz_problem* p, const double* U, int32_t objective, double scale,
                 double* work, double* J, double* dU, double* dG, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!U || !work || !J || !dU || !dG) return ofail(NZ_E_ARG, "null array");
    if (objective < NZ_OBJ_THRUST || objective > NZ_OBJ_EXIT_PRESSURE)
        return ofail(NZ_E_ARG, "unknown objective id");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nbq = (p->n_edge - p->n_int) * p->nqf;
    double* val = work;            // (nbq, 2)
    double* gS = work + 2 * nbq;   // (nbq, 4)
    double* red = gS + 4 * nbq;    // (2,)
    const bool averaged = objective == NZ_OBJ_EXIT_MACH || objective == NZ_OBJ_EXIT_PRESSURE;
    if (p->n_int > 0)
        hipLaunchKernelGGL(k_zero_interior_edges, dim3(nblk(p->n_int * p->nqf * 3)), dim3(BS), 0,
                           st, *p, dG);
    if (nbq > 0)
        hipLaunchKernelGGL(k_objective_points, dim3(nblk(nbq)), dim3(BS), 0, st, *p, U,
                           objective, scale, val, gS, dG);
    hipLaunchKernelGGL(k_objective_reduce, dim3(1), dim3(BS), 0, st, val, nbq, averaged ? 1 : 0,
                       red);
    if (averaged && nbq > 0)
        hipLaunchKernelGGL(k_objective_average, dim3(nblk(nbq)), dim3(BS), 0, st, *p, red, gS, dG);
    hipLaunchKernelGGL(k_objective_elements, dim3(nblk(p->n_elem)), dim3(BS), 0, st, *p, gS, dU);
    rc = ocheck("nz_objective");
    if (rc) return rc;
    if (hipMemcpyAsync(J, red, sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return ofail(NZ_E_LAUNCH, "nz_objective: copy of J");
    return NZ_OK;
}

}  // extern "C"
