// Pointwise Euler physics for the device kernels: primitives, the three
// interface fluxes and the boundary fluxes.  FP64 throughout.
//
// Formula forms follow the reference's NumPy path, which is the parity
// oracle: src/physics/thermo.py:46-86, src/physics/fluxes.py:38-445,
// src/physics/boundary.py:20-214.  Branches replace the reference's
// branchless `where` selections; every selected value is computed with the
// same operations, so the two agree to round-off.
#pragma once
#include <hip/hip_runtime.h>

#include "nozzle_hip.h"

namespace nz {

constexpr double kFloor = 1e-10;  // thermo.py:33

struct Prim {
    double r, u, v, p, H;
};

__device__ __forceinline__ Prim primitives(const double U[4], double g) {
    Prim w;
    w.r = fmax(U[0], kFloor);
    w.u = U[1] / w.r;
    w.v = U[2] / w.r;
    w.p = fmax((g - 1.0) * (U[3] - 0.5 * w.r * (w.u * w.u + w.v * w.v)), kFloor);
    w.H = (U[3] + w.p) / w.r;
    return w;
}

__device__ __forceinline__ void normal_flux(const Prim& w, double nx, double ny, double f[4]) {
    const double vn = w.u * nx + w.v * ny;
    f[0] = w.r * vn;
    f[1] = w.r * w.u * vn + w.p * nx;
    f[2] = w.r * w.v * vn + w.p * ny;
    f[3] = w.r * w.H * vn;
}

// Roe with the Harten-Hyman entropy fix (fluxes.py:38-103).
__device__ __forceinline__ double flux_roe(const double UL[4], const double UR[4], double nx,
                                           double ny, double g, double efix, double out[4]) {
    const Prim L = primitives(UL, g), R = primitives(UR, g);
    double FL[4], FR[4];
    normal_flux(L, nx, ny, FL);
    normal_flux(R, nx, ny, FR);
    const double sL = sqrt(L.r), sR = sqrt(R.r), den = sL + sR;
    const double vx = (sL * L.u + sR * R.u) / den;
    const double vy = (sL * L.v + sR * R.v) / den;
    const double H = (sL * L.H + sR * R.H) / den;
    const double q2 = vx * vx + vy * vy;
    const double c = sqrt(fmax((g - 1.0) * (H - 0.5 * q2), kFloor));
    const double vn = vx * nx + vy * ny;
    double l1 = fabs(vn + c), l2 = fabs(vn - c), l3 = fabs(vn);
    const double smax = fmax(l1, l2);
    const double eps = efix * c;
    if (l1 < eps) l1 = (l1 * l1 + eps * eps) / (2.0 * eps);
    if (l2 < eps) l2 = (l2 * l2 + eps * eps) / (2.0 * eps);
    if (l3 < eps) l3 = (l3 * l3 + eps * eps) / (2.0 * eps);
    const double d0 = UR[0] - UL[0], d1 = UR[1] - UL[1], d2 = UR[2] - UL[2], d3 = UR[3] - UL[3];
    const double G1 = (g - 1.0) * (0.5 * q2 * d0 - vx * d1 - vy * d2 + d3);
    const double G2 = -vn * d0 + d1 * nx + d2 * ny;
    const double s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2);
    const double C1 = (G1 / (c * c)) * (s1 - l3) + (G2 / c) * s2;
    const double C2 = (G1 / c) * s2 + (s1 - l3) * G2;
    out[0] = 0.5 * (FL[0] + FR[0]) - 0.5 * (l3 * d0 + C1);
    out[1] = 0.5 * (FL[1] + FR[1]) - 0.5 * (l3 * d1 + C1 * vx + C2 * nx);
    out[2] = 0.5 * (FL[2] + FR[2]) - 0.5 * (l3 * d2 + C1 * vy + C2 * ny);
    out[3] = 0.5 * (FL[3] + FR[3]) - 0.5 * (l3 * d3 + C1 * H + C2 * vn);
    return smax;
}

__device__ __forceinline__ double away_from_zero(double x) { return fabs(x) > kFloor ? x : kFloor; }

// HLLC with Batten's wave speeds and the optional low-Mach scaling (fluxes.py:106-255).
__device__ __forceinline__ double flux_hllc(const double UL[4], const double UR[4], double nx,
                                            double ny, double g, double low_mach, double out[4]) {
    const Prim L = primitives(UL, g), R = primitives(UR, g);
    const double vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    const double aL = sqrt(fmax(g * L.p / L.r, kFloor)), aR = sqrt(fmax(g * R.p / R.r, kFloor));
    const double sL = sqrt(L.r), sR = sqrt(R.r), den = sL + sR;
    const double vx = (sL * L.u + sR * R.u) / den;
    const double vy = (sL * L.v + sR * R.v) / den;
    const double H = (sL * L.H + sR * R.H) / den;
    const double vnt = vx * nx + vy * ny;
    const double at = sqrt(fmax((g - 1.0) * (H - 0.5 * (vx * vx + vy * vy)), kFloor));
    double SL = fmin(vnL - aL, vnt - at);
    double SR = fmax(vnR + aR, vnt + at);
    if (low_mach > 0.0) {
        const double mach = fmax(fabs(vnL) / aL, fabs(vnR) / aR);
        const double phi = fmin(1.0, mach / low_mach);
        const double lo = fmin(vnL, vnt), hi = fmax(vnR, vnt);
        SL = lo - phi * (lo - SL);
        SR = hi + phi * (SR - hi);
    }
    const double mL = L.r * (SL - vnL), mR = R.r * (SR - vnR);
    const double SM =
        (R.r * vnR * (SR - vnR) - L.r * vnL * (SL - vnL) + L.p - R.p) / away_from_zero(mR - mL);
    const double ms = fmax(fabs(SL), fabs(SR));
    // one side's flux and star state, selected before evaluation
    const bool left = (SL >= 0.0) || (SR > 0.0 && SM >= 0.0);
    const Prim& W = left ? L : R;
    const double* U = left ? UL : UR;
    double F[4];
    normal_flux(W, nx, ny, F);
    if (SL >= 0.0 || SR <= 0.0) {
        for (int s = 0; s < 4; ++s) out[s] = F[s];
        return ms;
    }
    const double S = left ? SL : SR, vn = left ? vnL : vnR;
    const double fac = (S - vn) / away_from_zero(S - SM);
    const double rs = W.r * fac;
    const double us = W.u + (SM - vn) * nx, vs = W.v + (SM - vn) * ny;
    const double Es = U[3] / W.r + (SM - vn) * (SM + W.p / away_from_zero(W.r * (S - vn)));
    out[0] = F[0] + S * (rs - U[0]);
    out[1] = F[1] + S * (rs * us - U[1]);
    out[2] = F[2] + S * (rs * vs - U[2]);
    out[3] = F[3] + S * (rs * Es - U[3]);
    return ms;
}

// AUSM+-up (fluxes.py:261-445): beta 1/8, alpha 3/16, Kp 0.4, Ku 0.2.
__device__ __forceinline__ double sq(double x) { return x * x; }

__device__ __forceinline__ double flux_ausm(const double UL[4], const double UR[4], double nx,
                                            double ny, double g, double out[4]) {
    const Prim L = primitives(UL, g), R = primitives(UR, g);
    const double vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    const double aL = sqrt(g * L.p / L.r), aR = sqrt(g * R.p / R.r);
    const double ah = 0.5 * (aL + aR), rh = 0.5 * (L.r + R.r);
    const double mL = vnL / ah, mR = vnR / ah;
    const double mb2 = 0.5 * (mL * mL + mR * mR);
    // split Mach numbers M4+(mL), M4-(mR)
    double m4p, m4m;
    if (fabs(mL) >= 1.0) m4p = 0.5 * (mL + fabs(mL));
    else m4p = (0.25 * sq(mL + 1.0)) * (1.0 - 16.0 * 0.125 * (-0.25 * sq(mL - 1.0)));
    if (fabs(mR) >= 1.0) m4m = 0.5 * (mR - fabs(mR));
    else m4m = (-0.25 * sq(mR - 1.0)) * (1.0 + 16.0 * 0.125 * (0.25 * sq(mR + 1.0)));
    const double uh = ah * (m4p + m4m) + 0.4 * fmax(1.0 - mb2, 0.0) * (L.p - R.p) / (rh * ah);
    // pressure splittings P5+(mL), P5-(mR); the supersonic branch is (1 +- sign M)/2
    double pp, pm;
    if (fabs(mL) >= 1.0) pp = mL > 0.0 ? 1.0 : 0.0;
    else pp = (0.25 * sq(mL + 1.0)) * ((2.0 - mL) - 16.0 * 0.1875 * mL * (-0.25 * sq(mL - 1.0)));
    if (fabs(mR) >= 1.0) pm = mR > 0.0 ? 0.0 : 1.0;
    else pm = (-0.25 * sq(mR - 1.0)) * ((-2.0 - mR) + 16.0 * 0.1875 * mR * (0.25 * sq(mR + 1.0)));
    const double ph = pp * L.p + pm * R.p + 0.2 * pp * pm * rh * ah * (vnL - vnR);
    const double PL[4] = {L.r, L.r * L.u, L.r * L.v, UL[3] + L.p};
    const double PR[4] = {R.r, R.r * R.u, R.r * R.v, UR[3] + R.p};
    const double au = fabs(uh);
    for (int s = 0; s < 4; ++s) out[s] = 0.5 * uh * (PL[s] + PR[s]) - 0.5 * au * (PR[s] - PL[s]);
    out[1] += ph * nx;
    out[2] += ph * ny;
    return fmax(fabs(vnL) + aL, fabs(vnR) + aR);
}

// Slip wall and symmetry axis (boundary.py:20-38).
__device__ __forceinline__ double flux_wall(const double U[4], double nx, double ny, double g,
                                            double out[4]) {
    const double r = fmax(U[0], kFloor), u = U[1] / r, v = U[2] / r;
    const double vn = u * nx + v * ny;
    const double tx = u - vn * nx, ty = v - vn * ny;
    const double pb = fmax((g - 1.0) * (U[3] - 0.5 * r * (tx * tx + ty * ty)), kFloor);
    out[0] = 0.0;
    out[1] = pb * nx;
    out[2] = pb * ny;
    out[3] = 0.0;
    return sqrt(g * pb / r);
}

// Subsonic stagnation inflow (boundary.py:41-110).
__device__ __forceinline__ double flux_inflow(const double U[4], double nx, double ny,
                                              const nz_problem& P, double out[4]) {
    const double g = P.gamma;
    const Prim w = primitives(U, g);
    const double a = sqrt(g * w.p / w.r);
    const double vn = w.u * nx + w.v * ny;
    const double Jp = vn + 2.0 * a / (g - 1.0);
    const double beta = (Jp / P.at) * (Jp / P.at);
    const double nd = nx * P.ca + ny * P.sa;
    double aa = 0.5 * (g - 1.0) * beta - nd * nd;
    const double bb = -4.0 * nd / (g - 1.0);
    const double t2 = 2.0 / (g - 1.0);
    const double cc = beta - t2 * t2;
    if (fabs(aa) < 1e-12) aa = aa >= 0.0 ? 1e-12 : -1e-12;
    const double root = sqrt(fmax(bb * bb - 4.0 * aa * cc, 0.0));
    const double r1 = (-bb + root) / (2.0 * aa), r2 = (-bb - root) / (2.0 * aa);
    double M = fmin(r1 >= 0.0 ? r1 : 1e30, r2 >= 0.0 ? r2 : 1e30);
    if (M > 0.5e30) M = 0.0;
    M = fmin(fmax(M, 0.0), 1.0);
    const double fac = 1.0 + 0.5 * (g - 1.0) * M * M;
    const double ab = sqrt(P.at2 / fac);
    const double qb = M * ab, vxb = qb * P.ca, vyb = qb * P.sa;
    const double rhob = P.rho_t * pow(fac, -1.0 / (g - 1.0));
    const double pb = rhob * ab * ab / g;
    const double Hb = P.at2 / (g - 1.0);
    const double vnb = vxb * nx + vyb * ny;
    out[0] = rhob * vnb;
    out[1] = rhob * vxb * vnb + pb * nx;
    out[2] = rhob * vyb * vnb + pb * ny;
    out[3] = rhob * Hb * vnb;
    return fabs(vnb) + ab;
}

// Pressure outflow with supersonic extrapolation and the optional backflow
// branch (boundary.py:113-214, including its max-speed definition).
__device__ __forceinline__ double flux_outflow(const double U[4], double nx, double ny,
                                               const nz_problem& P, double out[4]) {
    const double g = P.gamma;
    const Prim w = primitives(U, g);
    const double a = sqrt(g * w.p / w.r);
    const double vn = w.u * nx + w.v * ny;
    double Ub[4];
    if (vn / a >= 1.0) {
        for (int s = 0; s < 4; ++s) Ub[s] = U[s];
    } else {
        const double pb = P.p_back;
        const double rhob = w.r * pow(pb / w.p, 1.0 / g);
        const double ab = sqrt(g * pb / rhob);
        const double vnb = vn + 2.0 / (g - 1.0) * (a - ab);
        const double vxb = (w.u - vn * nx) + vnb * nx, vyb = (w.v - vn * ny) + vnb * ny;
        const double Eb = pb / ((g - 1.0) * rhob) + 0.5 * (vxb * vxb + vyb * vyb);
        Ub[0] = rhob;
        Ub[1] = rhob * vxb;
        Ub[2] = rhob * vyb;
        Ub[3] = rhob * Eb;
        if (P.rho_t_bf > 0.0) {
            const double rf = P.rho_t_bf * pow(pb / P.p_t_bf, 1.0 / g);
            const double af = sqrt(g * pb / rf);
            const double vnf = vn + 2.0 / (g - 1.0) * (a - af);
            const double vxf = vnf * nx, vyf = vnf * ny;
            const double Ef = pb / ((g - 1.0) * rf) + 0.5 * (vxf * vxf + vyf * vyf);
            const double Uf[4] = {rf, rf * vxf, rf * vyf, rf * Ef};
            const double t = fmin(fmax((vn / a) / fmax(P.band_bf, 1e-12), 0.0), 1.0);
            for (int s = 0; s < 4; ++s) Ub[s] = t * Ub[s] + (1.0 - t) * Uf[s];
        }
    }
    const Prim o = primitives(Ub, g);
    normal_flux(o, nx, ny, out);
    return fabs(o.u * nx + o.v * ny) + sqrt(g * o.p / o.r);
}

// FLUX: 0 roe, 1 hllc, 2 ausmup (src/backends/numba_backend.py:59-67)
template <int FLUX>
__device__ __forceinline__ double interface_flux(const nz_problem& P, const double UL[4],
                                                 const double UR[4], double nx, double ny,
                                                 double out[4]) {
    if constexpr (FLUX == 1) return flux_hllc(UL, UR, nx, ny, P.gamma, P.low_mach, out);
    else if constexpr (FLUX == 2) return flux_ausm(UL, UR, nx, ny, P.gamma, out);
    else return flux_roe(UL, UR, nx, ny, P.gamma, P.efix, out);
}

// tag: 1 inflow, 2 outflow, 3 wall, 4 axis (src/geometry/mesh.py:34-40)
__device__ __forceinline__ double boundary_flux(const nz_problem& P, int tag, const double U[4],
                                                double nx, double ny, double out[4]) {
    if (tag == 1) return flux_inflow(U, nx, ny, P, out);
    if (tag == 2) return flux_outflow(U, nx, ny, P, out);
    return flux_wall(U, nx, ny, P.gamma, out);
}

}  // namespace nz
