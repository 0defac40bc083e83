// Pointwise Euler physics for the device kernels: primitives, the three
// interface fluxes and the boundary fluxes.  FP64 throughout.
//
// Formula forms follow the reference's NumPy path, which is the parity
// oracle: src/physics/thermo.py:46-86, src/physics/fluxes.py:38-445,
// src/physics/boundary.py:20-214.  Branches replace the reference's
// branchless `where` selections; every selected value is computed with the
// same operations, so the two agree to round-off.
//
// Every function is a template over the scalar type: `double` in the
// forward kernels, forward-mode `Dual<K>` in the adjoint kernels, so the
// flux Jacobians are those of exactly the arithmetic evaluated here.  The
// interface fluxes take the normal as a type of its own: `double` when only
// the state is differentiated, a dual when the normal is as well.
#pragma once
#include <hip/hip_runtime.h>

#include "dual.hip.h"
#include "nozzle_hip.h"

namespace nz {

constexpr double kFloor = 1e-10;  // thermo.py:33

// Reciprocal and square root of the flux formulas' operands, which the
// floors keep positive and normal (densities and pressures >= 1e-10, sound
// speeds, sums of square roots of densities).  The hardware estimates
// (v_rcp_f64, v_rsq_f64) are refined by Newton-Raphson steps: for 1/x,
// r <- r + r (1 - x r) twice; for sqrt(x), one coupled step on g ~ sqrt(x),
// h ~ 1/(2 sqrt(x)) followed by two residual corrections g <- g + h (x - g^2).
// Each lands within about an ulp of the correctly rounded result in 5 and
// 9 instructions, against the IEEE sequences' 10 and 15, whose extra work
// handles denormal, zero and infinite operands that cannot occur here.
// NaN propagates.  The dual-number overloads differentiate 1/x and sqrt(x).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double root(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

template <int K>
__device__ __forceinline__ Dual<K> recip(const Dual<K>& x) {
    return 1.0 / x;
}

template <int K>
__device__ __forceinline__ Dual<K> root(const Dual<K>& x) {
    return sqrt(x);
}

template <class T>
struct Prim {
    T r, u, v, p, H, ir;  // ir = 1 / r
};

template <class T>
__device__ __forceinline__ Prim<T> primitives(const T U[4], double g) {
    // one reciprocal instead of three divisions; the results differ from
    // the quotients by about an ulp
    Prim<T> w;
    w.r = fmax(U[0], kFloor);
    const T ir = recip(w.r);
    w.u = U[1] * ir;
    w.v = U[2] * ir;
    w.p = fmax((g - 1.0) * (U[3] - 0.5 * w.r * (w.u * w.u + w.v * w.v)), kFloor);
    w.H = (U[3] + w.p) * ir;
    w.ir = ir;
    return w;
}

template <class T, class N_>
__device__ __forceinline__ void normal_flux(const Prim<T>& w, const N_& nx, const N_& ny, T f[4]) {
    const T vn = w.u * nx + w.v * ny;
    f[0] = w.r * vn;
    f[1] = w.r * w.u * vn + w.p * nx;
    f[2] = w.r * w.v * vn + w.p * ny;
    f[3] = w.r * w.H * vn;
}

// The boundary fluxes run in the boundary pre-pass and, for chained
// Runge-Kutta stages, in the stage kernels' epilogues.  With FMA contraction
// left to the compiler, a sum of two products can fuse either product
// depending on the code around it, so the two places would round apart.
// These forms and the boundary fluxes below are compiled without
// contraction: every product and sum rounds on its own, wherever inlined.
template <class T>
__device__ __forceinline__ Prim<T> primitives_x(const T U[4], double g) {
#pragma clang fp contract(off)
    Prim<T> w;
    w.r = fmax(U[0], kFloor);
    const T ir = recip(w.r);
    w.u = U[1] * ir;
    w.v = U[2] * ir;
    w.p = fmax((g - 1.0) * (U[3] - 0.5 * w.r * (w.u * w.u + w.v * w.v)), kFloor);
    w.H = (U[3] + w.p) * ir;
    w.ir = ir;
    return w;
}

template <class T, class N_>
__device__ __forceinline__ void normal_flux_x(const Prim<T>& w, const N_& nx, const N_& ny, T f[4]) {
#pragma clang fp contract(off)
    const T vn = w.u * nx + w.v * ny;
    f[0] = w.r * vn;
    f[1] = w.r * w.u * vn + w.p * nx;
    f[2] = w.r * w.v * vn + w.p * ny;
    f[3] = w.r * w.H * vn;
}

// Roe with the Harten-Hyman entropy fix (fluxes.py:38-103).
template <class T, class N_>
__device__ __forceinline__ T flux_roe(const T UL[4], const T UR[4], const N_& nx, const N_& ny,
                                      double g, double efix, T out[4]) {
    // Ordered so that few values are live at once: the central flux and the
    // jumps are formed first and the one-sided states die with the Roe means.
    // The three entropy-fix quotients share one reciprocal.
    const Prim<T> L = primitives(UL, g), R = primitives(UR, g);
    T FL[4], FR[4];
    normal_flux(L, nx, ny, FL);
    normal_flux(R, nx, ny, FR);
    T Fc[4], d[4];
    for (int k = 0; k < 4; ++k) {
        Fc[k] = 0.5 * (FL[k] + FR[k]);
        d[k] = UR[k] - UL[k];
    }
    const T sL = root(L.r), sR = root(R.r), iden = recip(sL + sR);
    const T vx = (sL * L.u + sR * R.u) * iden;
    const T vy = (sL * L.v + sR * R.v) * iden;
    const T H = (sL * L.H + sR * R.H) * iden;
    const T q2 = vx * vx + vy * vy;
    const T c = root(fmax((g - 1.0) * (H - 0.5 * q2), kFloor));
    const T ic = recip(c);
    const T vn = vx * nx + vy * ny;
    T l1 = fabs(vn + c), l2 = fabs(vn - c), l3 = fabs(vn);
    const T smax = fmax(l1, l2);
    const T eps = efix * c;
    const T h = 0.5 * recip(eps);
    if (l1 < eps) l1 = (l1 * l1 + eps * eps) * h;
    if (l2 < eps) l2 = (l2 * l2 + eps * eps) * h;
    if (l3 < eps) l3 = (l3 * l3 + eps * eps) * h;
    const T G1 = (g - 1.0) * (0.5 * q2 * d[0] - vx * d[1] - vy * d[2] + d[3]);
    const T G2 = -vn * d[0] + d[1] * nx + d[2] * ny;
    const T s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2);
    const T G1c = G1 * ic;
    const T C1 = (G1c * ic) * (s1 - l3) + (G2 * ic) * s2;
    const T C2 = G1c * s2 + (s1 - l3) * G2;
    out[0] = Fc[0] - 0.5 * (l3 * d[0] + C1);
    out[1] = Fc[1] - 0.5 * (l3 * d[1] + C1 * vx + C2 * nx);
    out[2] = Fc[2] - 0.5 * (l3 * d[2] + C1 * vy + C2 * ny);
    out[3] = Fc[3] - 0.5 * (l3 * d[3] + C1 * H + C2 * vn);
    return smax;
}

template <class T>
__device__ __forceinline__ T away_from_zero(const T& x) {
    return fabs(x) > kFloor ? x : T(kFloor);
}

// HLLC with Batten's wave speeds and the optional low-Mach scaling (fluxes.py:106-255).
template <class T, class N_>
__device__ __forceinline__ T flux_hllc(const T UL[4], const T UR[4], const N_& nx, const N_& ny,
                                       double g, double low_mach, T out[4]) {
    const Prim<T> L = primitives(UL, g), R = primitives(UR, g);
    const T vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    const T aL = root(fmax(g * L.p * L.ir, kFloor)), aR = root(fmax(g * R.p * R.ir, kFloor));
    const T sL = root(L.r), sR = root(R.r), iden = recip(sL + sR);
    const T vx = (sL * L.u + sR * R.u) * iden;
    const T vy = (sL * L.v + sR * R.v) * iden;
    const T H = (sL * L.H + sR * R.H) * iden;
    const T vnt = vx * nx + vy * ny;
    const T at = root(fmax((g - 1.0) * (H - 0.5 * (vx * vx + vy * vy)), kFloor));
    T SL = fmin(vnL - aL, vnt - at);
    T SR = fmax(vnR + aR, vnt + at);
    if (low_mach > 0.0) {
        const T mach = fmax(fabs(vnL) / aL, fabs(vnR) / aR);
        const T phi = fmin(T(1.0), mach / low_mach);
        const T lo = fmin(vnL, vnt), hi = fmax(vnR, vnt);
        SL = lo - phi * (lo - SL);
        SR = hi + phi * (SR - hi);
    }
    const T mL = L.r * (SL - vnL), mR = R.r * (SR - vnR);
    const T SM = (R.r * vnR * (SR - vnR) - L.r * vnL * (SL - vnL) + L.p - R.p) / away_from_zero(mR - mL);
    const T ms = fmax(fabs(SL), fabs(SR));
    // one side's flux and star state, selected before evaluation
    const bool left = (SL >= 0.0) || (SR > 0.0 && SM >= 0.0);
    const Prim<T>& W = left ? L : R;
    const T* U = left ? UL : UR;
    T F[4];
    normal_flux(W, nx, ny, F);
    if (SL >= 0.0 || SR <= 0.0) {
        for (int s = 0; s < 4; ++s) out[s] = F[s];
        return ms;
    }
    const T S = left ? SL : SR, vn = left ? vnL : vnR;
    const T fac = (S - vn) / away_from_zero(S - SM);
    const T rs = W.r * fac;
    const T us = W.u + (SM - vn) * nx, vs = W.v + (SM - vn) * ny;
    const T Es = U[3] / W.r + (SM - vn) * (SM + W.p / away_from_zero(W.r * (S - vn)));
    out[0] = F[0] + S * (rs - U[0]);
    out[1] = F[1] + S * (rs * us - U[1]);
    out[2] = F[2] + S * (rs * vs - U[2]);
    out[3] = F[3] + S * (rs * Es - U[3]);
    return ms;
}

// AUSM+-up (fluxes.py:261-445): beta 1/8, alpha 3/16, Kp 0.4, Ku 0.2.
template <class T>
__device__ __forceinline__ T sq(const T& x) {
    return x * x;
}

template <class T, class N_>
__device__ __forceinline__ T flux_ausm(const T UL[4], const T UR[4], const N_& nx, const N_& ny,
                                       double g, T out[4]) {
    const Prim<T> L = primitives(UL, g), R = primitives(UR, g);
    const T vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    const T aL = root(g * L.p * L.ir), aR = root(g * R.p * R.ir);
    const T ah = 0.5 * (aL + aR), rh = 0.5 * (L.r + R.r);
    const T iah = recip(ah);
    const T mL = vnL * iah, mR = vnR * iah;
    const T mb2 = 0.5 * (mL * mL + mR * mR);
    // split Mach numbers M4+(mL), M4-(mR)
    T m4p, m4m;
    if (fabs(mL) >= 1.0) m4p = 0.5 * (mL + fabs(mL));
    else m4p = (0.25 * sq(mL + 1.0)) * (1.0 - 16.0 * 0.125 * (-0.25 * sq(mL - 1.0)));
    if (fabs(mR) >= 1.0) m4m = 0.5 * (mR - fabs(mR));
    else m4m = (-0.25 * sq(mR - 1.0)) * (1.0 + 16.0 * 0.125 * (0.25 * sq(mR + 1.0)));
    const T uh = ah * (m4p + m4m) + 0.4 * fmax(1.0 - mb2, 0.0) * (L.p - R.p) * recip(rh * ah);
    // pressure splittings P5+(mL), P5-(mR); the supersonic branch is (1 +- sign M)/2
    T pp, pm;
    if (fabs(mL) >= 1.0) pp = T(mL > 0.0 ? 1.0 : 0.0);
    else pp = (0.25 * sq(mL + 1.0)) * ((2.0 - mL) - 16.0 * 0.1875 * mL * (-0.25 * sq(mL - 1.0)));
    if (fabs(mR) >= 1.0) pm = T(mR > 0.0 ? 0.0 : 1.0);
    else pm = (-0.25 * sq(mR - 1.0)) * ((-2.0 - mR) + 16.0 * 0.1875 * mR * (0.25 * sq(mR + 1.0)));
    const T ph = pp * L.p + pm * R.p + 0.2 * pp * pm * rh * ah * (vnL - vnR);
    const T PL[4] = {L.r, L.r * L.u, L.r * L.v, UL[3] + L.p};
    const T PR[4] = {R.r, R.r * R.u, R.r * R.v, UR[3] + R.p};
    const T au = fabs(uh);
    for (int s = 0; s < 4; ++s) out[s] = 0.5 * uh * (PL[s] + PR[s]) - 0.5 * au * (PR[s] - PL[s]);
    out[1] += ph * nx;
    out[2] += ph * ny;
    return fmax(fabs(vnL) + aL, fabs(vnR) + aR);
}

// Slip wall and symmetry axis (boundary.py:20-38).
template <class T>
__device__ __forceinline__ T flux_wall(const T U[4], const T& nx, const T& ny, double g,
                                       T out[4]) {
#pragma clang fp contract(off)
    const T r = fmax(U[0], kFloor), u = U[1] / r, v = U[2] / r;
    const T vn = u * nx + v * ny;
    const T tx = u - vn * nx, ty = v - vn * ny;
    const T pb = fmax((g - 1.0) * (U[3] - 0.5 * r * (tx * tx + ty * ty)), kFloor);
    out[0] = T(0.0);
    out[1] = pb * nx;
    out[2] = pb * ny;
    out[3] = T(0.0);
    return sqrt(g * pb / r);
}

// Subsonic stagnation inflow (boundary.py:41-110).
template <class T>
__device__ __forceinline__ T flux_inflow(const T U[4], const T& nx, const T& ny, const nz_problem& P,
                                         T out[4]) {
#pragma clang fp contract(off)
    const double g = P.gamma;
    const Prim<T> w = primitives_x(U, g);
    const T a = sqrt(g * w.p / w.r);
    const T vn = w.u * nx + w.v * ny;
    const T Jp = vn + 2.0 * a / (g - 1.0);
    const T beta = (Jp / P.at) * (Jp / P.at);
    const T nd = nx * P.ca + ny * P.sa;
    T aa = 0.5 * (g - 1.0) * beta - nd * nd;
    const T bb = -4.0 * nd / (g - 1.0);
    const double t2 = 2.0 / (g - 1.0);
    const T cc = beta - t2 * t2;
    if (fabs(aa) < 1e-12) aa = T(aa >= 0.0 ? 1e-12 : -1e-12);
    const T root = sqrt(fmax(bb * bb - 4.0 * aa * cc, 0.0));
    const T r1 = (-bb + root) / (2.0 * aa), r2 = (-bb - root) / (2.0 * aa);
    T M = fmin(r1 >= 0.0 ? r1 : T(1e30), r2 >= 0.0 ? r2 : T(1e30));
    if (M > 0.5e30) M = T(0.0);
    M = fmin(fmax(M, 0.0), 1.0);
    const T fac = 1.0 + 0.5 * (g - 1.0) * M * M;
    const T ab = sqrt(P.at2 / fac);
    const T qb = M * ab, vxb = qb * P.ca, vyb = qb * P.sa;
    const T rhob = P.rho_t * pow(fac, -1.0 / (g - 1.0));
    const T pb = rhob * ab * ab / g;
    const double Hb = P.at2 / (g - 1.0);
    const T vnb = vxb * nx + vyb * ny;
    out[0] = rhob * vnb;
    out[1] = rhob * vxb * vnb + pb * nx;
    out[2] = rhob * vyb * vnb + pb * ny;
    out[3] = rhob * Hb * vnb;
    return fabs(vnb) + ab;
}

// Pressure outflow with supersonic extrapolation and the optional backflow
// branch (boundary.py:113-214, including its max-speed definition).  The
// back pressure is an argument so the adjoint can differentiate through it.
template <class T>
__device__ __forceinline__ T flux_outflow(const T U[4], const T& nx, const T& ny,
                                          const nz_problem& P, const T& pb, T out[4]) {
#pragma clang fp contract(off)
    const double g = P.gamma;
    const Prim<T> w = primitives_x(U, g);
    const T a = sqrt(g * w.p / w.r);
    const T vn = w.u * nx + w.v * ny;
    T Ub[4];
    if (vn / a >= 1.0) {
        for (int s = 0; s < 4; ++s) Ub[s] = U[s];
    } else {
        const T rhob = w.r * pow(pb / w.p, 1.0 / g);
        const T ab = sqrt(g * pb / rhob);
        const T vnb = vn + 2.0 / (g - 1.0) * (a - ab);
        const T vxb = (w.u - vn * nx) + vnb * nx, vyb = (w.v - vn * ny) + vnb * ny;
        const T Eb = pb / ((g - 1.0) * rhob) + 0.5 * (vxb * vxb + vyb * vyb);
        Ub[0] = rhob;
        Ub[1] = rhob * vxb;
        Ub[2] = rhob * vyb;
        Ub[3] = rhob * Eb;
        if (P.rho_t_bf > 0.0) {
            const T rf = P.rho_t_bf * pow(pb / P.p_t_bf, 1.0 / g);
            const T af = sqrt(g * pb / rf);
            const T vnf = vn + 2.0 / (g - 1.0) * (a - af);
            const T vxf = vnf * nx, vyf = vnf * ny;
            const T Ef = pb / ((g - 1.0) * rf) + 0.5 * (vxf * vxf + vyf * vyf);
            const T Uf[4] = {rf, rf * vxf, rf * vyf, rf * Ef};
            const T t = fmin(fmax((vn / a) / fmax(P.band_bf, 1e-12), 0.0), 1.0);
            for (int s = 0; s < 4; ++s) Ub[s] = t * Ub[s] + (1.0 - t) * Uf[s];
        }
    }
    const Prim<T> o = primitives_x(Ub, g);
    normal_flux_x(o, nx, ny, out);
    return fabs(o.u * nx + o.v * ny) + sqrt(g * o.p / o.r);
}

// Reverse mode of the Euler fluxes at a volume point: with F(U), G(U) as
// evaluated by primitives() and the point-flux formulas, returns
// ub = (dF/dU)^T fb + (dG/dU)^T gb.  Each statement of the forward evaluation
// is adjointed in reverse order, so the result is the derivative of exactly
// that arithmetic, floors included: a floored density or pressure passes no
// derivative (the taken branch of fmax).
__device__ __forceinline__ void euler_flux_vjp(const double U[4], double g, const double fb[4],
                                               const double gb[4], double ub[4]) {
    // forward
    const bool rok = U[0] >= kFloor;
    const double r = rok ? U[0] : kFloor;
    const double ir = recip(r);
    const double u = U[1] * ir, v = U[2] * ir;
    const double k = u * u + v * v;
    const double pq = (g - 1.0) * (U[3] - 0.5 * r * k);
    const bool pok = pq >= kFloor;
    const double p = pok ? pq : kFloor;
    const double H = (U[3] + p) * ir;
    // F = (r u, r u u + p, r u v, r u H), G = (r v, r u v, r v v + p, r v H)
    double rb = fb[0] * u + fb[1] * u * u + fb[2] * u * v + fb[3] * u * H
              + gb[0] * v + gb[1] * u * v + gb[2] * v * v + gb[3] * v * H;
    double ubar = fb[0] * r + fb[1] * 2.0 * r * u + fb[2] * r * v + fb[3] * r * H + gb[1] * r * v;
    double vbar = fb[2] * r * u + gb[0] * r + gb[1] * r * u + gb[2] * 2.0 * r * v + gb[3] * r * H;
    double pb = fb[1] + gb[2];
    const double Hb = fb[3] * r * u + gb[3] * r * v;
    // H = (U3 + p) ir
    double u3b = Hb * ir;
    pb += Hb * ir;
    double irb = Hb * (U[3] + p);
    // p = max((g - 1)(U3 - r k / 2), floor)
    if (pok) {
        const double qb = (g - 1.0) * pb;
        u3b += qb;
        rb -= 0.5 * k * qb;
        ubar -= r * u * qb;
        vbar -= r * v * qb;
    }
    // u = U1 ir, v = U2 ir
    const double u1b = ubar * ir, u2b = vbar * ir;
    irb += ubar * U[1] + vbar * U[2];
    // ir = 1 / r
    rb -= irb * ir * ir;
    ub[0] = rok ? rb : 0.0;
    ub[1] = u1b;
    ub[2] = u2b;
    ub[3] = u3b;
}

// Reverse mode of primitives(): the cotangents (rb, ub, vb, pb, Hb) of the
// primitive values pulled back to the conserved state U.
__device__ __forceinline__ void primitives_vjp(const double U[4], double g, double rb, double ub,
                                               double vb, double pb, double Hb, double out[4]) {
    const bool rok = U[0] >= kFloor;
    const double r = rok ? U[0] : kFloor;
    const double ir = recip(r);
    const double u = U[1] * ir, v = U[2] * ir;
    const double k = u * u + v * v;
    const double pq = (g - 1.0) * (U[3] - 0.5 * r * k);
    const double p = pq >= kFloor ? pq : kFloor;
    double u3b = Hb * ir;
    pb += Hb * ir;
    double irb = Hb * (U[3] + p);
    if (pq >= kFloor) {
        const double qb = (g - 1.0) * pb;
        u3b += qb;
        rb -= 0.5 * k * qb;
        ub -= r * u * qb;
        vb -= r * v * qb;
    }
    irb += ub * U[1] + vb * U[2];
    rb -= irb * ir * ir;
    out[0] = rok ? rb : 0.0;
    out[1] = ub * ir;
    out[2] = vb * ir;
    out[3] = u3b;
}

// Reverse mode of flux_roe with respect to one side's state (the left one
// when own_left): given the cotangent ob of the flux, returns
// ob^T dF/dU_own.  The statements of flux_roe are adjointed in reverse
// order, branch by branch as evaluated (entropy fixes, floors, |.|), so the
// result is the derivative of exactly that arithmetic, in one backward pass
// rather than four forward tangents.
__device__ __forceinline__ void roe_vjp(const double UL[4], const double UR[4], double nx,
                                        double ny, double g, double efix, bool own_left,
                                        const double ob[4], double out[4]) {
    // ---- forward, keeping what the backward pass reads
    const Prim<double> L = primitives(UL, g), R = primitives(UR, g);
    const double vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    double d[4];
    for (int k = 0; k < 4; ++k) d[k] = UR[k] - UL[k];
    const double sL = root(L.r), sR = root(R.r), iden = recip(sL + sR);
    const double nvx = sL * L.u + sR * R.u, nvy = sL * L.v + sR * R.v, nH = sL * L.H + sR * R.H;
    const double vx = nvx * iden, vy = nvy * iden, H = nH * iden;
    const double q2 = vx * vx + vy * vy;
    const double cq = (g - 1.0) * (H - 0.5 * q2);
    const bool cok = cq >= kFloor;
    const double c = root(cok ? cq : kFloor);
    const double ic = recip(c);
    const double vn = vx * nx + vy * ny;
    const double a1 = vn + c, a2 = vn - c;
    const double l1o = fabs(a1), l2o = fabs(a2), l3o = fabs(vn);
    const double eps = efix * c;
    const double re = recip(eps);
    const double h = 0.5 * re;
    const bool f1 = l1o < eps, f2 = l2o < eps, f3 = l3o < eps;
    const double l1 = f1 ? (l1o * l1o + eps * eps) * h : l1o;
    const double l2 = f2 ? (l2o * l2o + eps * eps) * h : l2o;
    const double l3 = f3 ? (l3o * l3o + eps * eps) * h : l3o;
    const double G1 = (g - 1.0) * (0.5 * q2 * d[0] - vx * d[1] - vy * d[2] + d[3]);
    const double G2 = -vn * d[0] + d[1] * nx + d[2] * ny;
    const double s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2);
    const double G1c = G1 * ic;
    const double t1 = G1c * ic, t2 = G2 * ic, sm = s1 - l3;
    const double C1 = t1 * sm + t2 * s2;
    const double C2 = G1c * s2 + sm * G2;
    // ---- backward
    double db[4];
    double l3b = -0.5 * (ob[0] * d[0] + ob[1] * d[1] + ob[2] * d[2] + ob[3] * d[3]);
    for (int k = 0; k < 4; ++k) db[k] = -0.5 * l3 * ob[k];
    const double C1b = -0.5 * (ob[0] + ob[1] * vx + ob[2] * vy + ob[3] * H);
    const double C2b = -0.5 * (ob[1] * nx + ob[2] * ny + ob[3] * vn);
    double vxb = -0.5 * ob[1] * C1, vyb = -0.5 * ob[2] * C1, Hb = -0.5 * ob[3] * C1;
    double vnb = -0.5 * ob[3] * C2;
    // C2 = G1c s2 + sm G2
    double G1cb = C2b * s2, s2b = C2b * G1c, smb = C2b * G2, G2b = C2b * sm;
    // C1 = t1 sm + t2 s2
    const double t1b = C1b * sm, t2b = C1b * s2;
    smb += C1b * t1;
    s2b += C1b * t2;
    G1cb += t1b * ic;
    double icb = t1b * G1c + t2b * G2;
    G2b += t2b * ic;
    // G1c = G1 ic
    const double G1b = G1cb * ic;
    icb += G1cb * G1;
    // sm = s1 - l3; s1 = (l1 + l2) / 2; s2 = (l1 - l2) / 2
    const double s1b = smb;
    l3b -= smb;
    double l1b = 0.5 * (s1b + s2b), l2b = 0.5 * (s1b - s2b);
    // G2 = -vn d0 + d1 nx + d2 ny
    vnb -= G2b * d[0];
    db[0] -= G2b * vn;
    db[1] += G2b * nx;
    db[2] += G2b * ny;
    // G1 = (g - 1)(q2 d0 / 2 - vx d1 - vy d2 + d3)
    const double gG = (g - 1.0) * G1b;
    double q2b = 0.5 * gG * d[0];
    db[0] += 0.5 * gG * q2;
    vxb -= gG * d[1];
    db[1] -= gG * vx;
    vyb -= gG * d[2];
    db[2] -= gG * vy;
    db[3] += gG;
    // entropy fixes l = (lo^2 + eps^2) h
    double epsb = 0.0, hb = 0.0;
    if (f1) {
        epsb += l1b * 2.0 * eps * h;
        hb += l1b * (l1o * l1o + eps * eps);
        l1b = l1b * 2.0 * l1o * h;
    }
    if (f2) {
        epsb += l2b * 2.0 * eps * h;
        hb += l2b * (l2o * l2o + eps * eps);
        l2b = l2b * 2.0 * l2o * h;
    }
    if (f3) {
        epsb += l3b * 2.0 * eps * h;
        hb += l3b * (l3o * l3o + eps * eps);
        l3b = l3b * 2.0 * l3o * h;
    }
    // h = re / 2, re = 1 / eps, eps = efix c
    epsb -= 0.5 * hb * re * re;
    double cb = epsb * efix;
    // l1 = |vn + c|, l2 = |vn - c|, l3 = |vn|
    const double a1b = a1 < 0.0 ? -l1b : l1b, a2b = a2 < 0.0 ? -l2b : l2b;
    vnb += a1b + a2b + (vn < 0.0 ? -l3b : l3b);
    cb += a1b - a2b;
    // vn = vx nx + vy ny
    vxb += vnb * nx;
    vyb += vnb * ny;
    // ic = 1 / c
    cb -= icb * ic * ic;
    // c = sqrt(max(cq, floor))
    if (cok) {
        const double cqb = cb * 0.5 * ic;
        Hb += (g - 1.0) * cqb;
        q2b -= 0.5 * (g - 1.0) * cqb;
    }
    // q2 = vx^2 + vy^2
    vxb += 2.0 * vx * q2b;
    vyb += 2.0 * vy * q2b;
    // vx = nvx iden, vy = nvy iden, H = nH iden
    const double nvxb = vxb * iden, nvyb = vyb * iden, nHb = Hb * iden;
    const double idenb = vxb * nvx + vyb * nvy + Hb * nH;
    // iden = 1 / (sL + sR)
    const double sumb = -idenb * iden * iden;
    // the own side only from here
    const Prim<double>& W = own_left ? L : R;
    const double sW = own_left ? sL : sR, vnW = own_left ? vnL : vnR;
    double sWb = sumb + nvxb * W.u + nvyb * W.v + nHb * W.H;
    double rb = sWb * 0.5 / sW;  // s = sqrt(r)
    double ub = nvxb * sW, vb = nvyb * sW, Wb = nHb * sW;
    // the central part (F_own / 2) of the flux: normal_flux(W)
    double fb[4];
    for (int k = 0; k < 4; ++k) fb[k] = 0.5 * ob[k];
    rb += fb[0] * vnW + fb[1] * W.u * vnW + fb[2] * W.v * vnW + fb[3] * W.H * vnW;
    ub += fb[1] * W.r * vnW;
    vb += fb[2] * W.r * vnW;
    Wb += fb[3] * W.r * vnW;
    const double pb = fb[1] * nx + fb[2] * ny;
    const double vnWb = fb[0] * W.r + fb[1] * W.r * W.u + fb[2] * W.r * W.v + fb[3] * W.r * W.H;
    ub += vnWb * nx;
    vb += vnWb * ny;
    primitives_vjp(own_left ? UL : UR, g, rb, ub, vb, pb, Wb, out);
    // d = UR - UL
    for (int k = 0; k < 4; ++k) out[k] += own_left ? -db[k] : db[k];
}

// Reverse mode of flux_ausm with respect to one side's state (the left one
// when own_left): returns ob^T dF/dU_own.  Every statement of flux_ausm is
// adjointed for both sides (the branches on |M| and on the sign of the
// interface velocity as evaluated; the max as the dual numbers take it, the
// first argument on ties), then the own side's primitive cotangents are
// pulled back to its conserved state.
__device__ __forceinline__ void ausm_vjp(const double UL[4], const double UR[4], double nx,
                                         double ny, double g, bool own_left, const double ob[4],
                                         double out[4]) {
    // ---- forward, keeping what the backward pass reads
    const Prim<double> L = primitives(UL, g), R = primitives(UR, g);
    const double vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    const double aL = root(g * L.p * L.ir), aR = root(g * R.p * R.ir);
    const double ah = 0.5 * (aL + aR), rh = 0.5 * (L.r + R.r);
    const double iah = recip(ah);
    const double mL = vnL * iah, mR = vnR * iah;
    const double mb2 = 0.5 * (mL * mL + mR * mR);
    const bool supL = fabs(mL) >= 1.0, supR = fabs(mR) >= 1.0;
    double m4p, m4m, pp, pm;
    // subsonic pieces: m4p = E F, m4m = G H, pp = A B, pm = C D
    const double E = 0.25 * (mL + 1.0) * (mL + 1.0), F = 1.0 + 0.5 * (mL - 1.0) * (mL - 1.0);
    const double G = -0.25 * (mR - 1.0) * (mR - 1.0), H = 1.0 + 0.5 * (mR + 1.0) * (mR + 1.0);
    const double B = (2.0 - mL) + 0.75 * mL * (mL - 1.0) * (mL - 1.0);
    const double D = (-2.0 - mR) + 0.75 * mR * (mR + 1.0) * (mR + 1.0);
    if (supL) m4p = 0.5 * (mL + fabs(mL));
    else m4p = (0.25 * sq(mL + 1.0)) * (1.0 - 16.0 * 0.125 * (-0.25 * sq(mL - 1.0)));
    if (supR) m4m = 0.5 * (mR - fabs(mR));
    else m4m = (-0.25 * sq(mR - 1.0)) * (1.0 + 16.0 * 0.125 * (0.25 * sq(mR + 1.0)));
    const double om = 1.0 - mb2;
    const double f = om >= 0.0 ? om : 0.0;
    const double irh = recip(rh * ah);
    const double dp = L.p - R.p;
    const double uh = ah * (m4p + m4m) + 0.4 * f * dp * irh;
    if (supL) pp = mL > 0.0 ? 1.0 : 0.0;
    else pp = (0.25 * sq(mL + 1.0)) * ((2.0 - mL) - 16.0 * 0.1875 * mL * (-0.25 * sq(mL - 1.0)));
    if (supR) pm = mR > 0.0 ? 0.0 : 1.0;
    else pm = (-0.25 * sq(mR - 1.0)) * ((-2.0 - mR) + 16.0 * 0.1875 * mR * (0.25 * sq(mR + 1.0)));
    const double dv = vnL - vnR;
    const double PL[4] = {L.r, L.r * L.u, L.r * L.v, UL[3] + L.p};
    const double PR[4] = {R.r, R.r * R.u, R.r * R.v, UR[3] + R.p};
    const double au = fabs(uh);
    // ---- backward
    double uhb = 0.0, aub = 0.0, PLb[4], PRb[4];
    for (int k = 0; k < 4; ++k) {
        uhb += 0.5 * ob[k] * (PL[k] + PR[k]);
        aub -= 0.5 * ob[k] * (PR[k] - PL[k]);
        PLb[k] = 0.5 * ob[k] * (uh + au);
        PRb[k] = 0.5 * ob[k] * (uh - au);
    }
    const double phb = ob[1] * nx + ob[2] * ny;
    uhb += uh < 0.0 ? -aub : aub;
    // ph = pp pL + pm pR + 0.2 pp pm rh ah (vnL - vnR)
    double Lpb = phb * pp, Rpb = phb * pm;
    const double ppb = phb * (L.p + 0.2 * pm * rh * ah * dv);
    const double pmb = phb * (R.p + 0.2 * pp * rh * ah * dv);
    double rhb = phb * 0.2 * pp * pm * ah * dv;
    double ahb = phb * 0.2 * pp * pm * rh * dv;
    double vnLb = phb * 0.2 * pp * pm * rh * ah, vnRb = -vnLb;
    // uh = ah (m4p + m4m) + 0.4 f (pL - pR) / (rh ah)
    ahb += uhb * (m4p + m4m);
    const double m4pb = uhb * ah, m4mb = uhb * ah;
    const double mb2b = om >= 0.0 ? -uhb * 0.4 * dp * irh : 0.0;
    Lpb += uhb * 0.4 * f * irh;
    Rpb -= uhb * 0.4 * f * irh;
    const double irhb = uhb * 0.4 * f * dp;
    rhb -= irhb * irh * irh * ah;
    ahb -= irhb * irh * irh * rh;
    // the split Mach numbers and pressures
    double mLb = 0.0, mRb = 0.0;
    if (supL) {
        mLb += m4pb * (mL > 0.0 ? 1.0 : 0.0);
    } else {
        mLb += m4pb * (0.5 * (mL + 1.0) * F + E * (mL - 1.0));
        mLb += ppb * (0.5 * (mL + 1.0) * B +
                      E * (-1.0 + 0.75 * (mL - 1.0) * (mL - 1.0) + 1.5 * mL * (mL - 1.0)));
    }
    if (supR) {
        mRb += m4mb * (mR < 0.0 ? 1.0 : 0.0);
    } else {
        mRb += m4mb * (-0.5 * (mR - 1.0) * H + G * (mR + 1.0));
        mRb += pmb * (-0.5 * (mR - 1.0) * D +
                      G * (-1.0 + 0.75 * (mR + 1.0) * (mR + 1.0) + 1.5 * mR * (mR + 1.0)));
    }
    mLb += mb2b * mL;
    mRb += mb2b * mR;
    vnLb += mLb * iah;
    vnRb += mRb * iah;
    const double iahb = mLb * vnL + mRb * vnR;
    ahb -= iahb * iah * iah;
    const double aLb = 0.5 * ahb, aRb = 0.5 * ahb;
    // the own side from here: a = sqrt(g p ir), vn = u nx + v ny, P
    const Prim<double>& W = own_left ? L : R;
    const double* Pb = own_left ? PLb : PRb;
    const double aW = own_left ? aL : aR, aWb = own_left ? aLb : aRb;
    const double vnWb = own_left ? vnLb : vnRb;
    double pb = own_left ? Lpb : Rpb;
    const double zb = aWb * 0.5 / aW;
    pb += zb * g * W.ir + Pb[3];
    const double irb = zb * g * W.p;
    double rb = 0.5 * rhb - irb * W.ir * W.ir + Pb[0] + Pb[1] * W.u + Pb[2] * W.v;
    const double ub = vnWb * nx + Pb[1] * W.r, vb = vnWb * ny + Pb[2] * W.r;
    primitives_vjp(own_left ? UL : UR, g, rb, ub, vb, pb, 0.0, out);
    out[3] += Pb[3];
}

// Reverse mode of flux_hllc with respect to one side's state (the left one
// when own_left): returns ob^T dF/dU_own.  The statements of flux_hllc are
// adjointed for both sides with the branches as evaluated (the wave-speed
// min/max and the low-Mach blend taking the selected argument, the first on
// ties, as the dual numbers do; the floors and away_from_zero passing no
// derivative when active), then the own side's primitive cotangents are
// pulled back to its conserved state.
__device__ __forceinline__ void hllc_vjp(const double UL[4], const double UR[4], double nx,
                                         double ny, double g, double low_mach, bool own_left,
                                         const double ob[4], double out[4]) {
    // ---- forward, keeping what the backward pass reads
    const Prim<double> L = primitives(UL, g), R = primitives(UR, g);
    const double vnL = L.u * nx + L.v * ny, vnR = R.u * nx + R.v * ny;
    const double zL = g * L.p * L.ir, zR = g * R.p * R.ir;
    const bool zLok = zL >= kFloor, zRok = zR >= kFloor;
    const double aL = root(zLok ? zL : kFloor), aR = root(zRok ? zR : kFloor);
    const double sL = root(L.r), sR = root(R.r), iden = recip(sL + sR);
    const double nvx = sL * L.u + sR * R.u, nvy = sL * L.v + sR * R.v, nH = sL * L.H + sR * R.H;
    const double vx = nvx * iden, vy = nvy * iden, H = nH * iden;
    const double vnt = vx * nx + vy * ny;
    const double cq = (g - 1.0) * (H - 0.5 * (vx * vx + vy * vy));
    const bool cok = cq >= kFloor;
    const double at = root(cok ? cq : kFloor);
    const double sl1 = vnL - aL, sl2 = vnt - at, sr1 = vnR + aR, sr2 = vnt + at;
    const bool slA = sl1 <= sl2, srA = sr1 >= sr2;
    const double SL0 = slA ? sl1 : sl2, SR0 = srA ? sr1 : sr2;
    double SL = SL0, SR = SR0;
    const bool lm = low_mach > 0.0;
    double phi = 1.0, lo = 0.0, hi = 0.0, xL = 0.0, xR = 0.0;
    bool phiA = true, machA = true, loA = true, hiA = true;
    if (lm) {
        xL = fabs(vnL) / aL;
        xR = fabs(vnR) / aR;
        machA = xL >= xR;
        const double mach = machA ? xL : xR;
        const double x = mach / low_mach;
        phiA = 1.0 <= x;
        phi = phiA ? 1.0 : x;
        loA = vnL <= vnt;
        hiA = vnR >= vnt;
        lo = loA ? vnL : vnt;
        hi = hiA ? vnR : vnt;
        SL = lo - phi * (lo - SL0);
        SR = hi + phi * (SR0 - hi);
    }
    const double mL = L.r * (SL - vnL), mR = R.r * (SR - vnR);
    const double num = R.r * vnR * (SR - vnR) - L.r * vnL * (SL - vnL) + L.p - R.p;
    const double dmm = mR - mL;
    const bool dmok = fabs(dmm) > kFloor;
    const double den = dmok ? dmm : kFloor;
    const double SM = num / den;
    const bool left = (SL >= 0.0) || (SR > 0.0 && SM >= 0.0);
    const Prim<double>& W = left ? L : R;
    const double* U = left ? UL : UR;
    const double vnW = left ? vnL : vnR;
    const bool star = !(SL >= 0.0 || SR <= 0.0);
    // cotangents of both sides' primitives (r, u, v, p, H, ir), of the selected
    // side's conserved state (direct), and of the shared intermediates
    double Lb[6] = {0, 0, 0, 0, 0, 0}, Rb[6] = {0, 0, 0, 0, 0, 0}, Ub[4] = {0, 0, 0, 0};
    double* Wb = left ? Lb : Rb;
    // ---- backward: out = F(W) [+ S (X - U)]
    {
        const double* fb = ob;  // normal_flux(W)
        Wb[0] += fb[0] * vnW + fb[1] * W.u * vnW + fb[2] * W.v * vnW + fb[3] * W.H * vnW;
        Wb[1] += fb[1] * W.r * vnW;
        Wb[2] += fb[2] * W.r * vnW;
        Wb[4] += fb[3] * W.r * vnW;
        Wb[3] += fb[1] * nx + fb[2] * ny;
        const double vb = fb[0] * W.r + fb[1] * W.r * W.u + fb[2] * W.r * W.v + fb[3] * W.r * W.H;
        Wb[1] += vb * nx;
        Wb[2] += vb * ny;
    }
    double SLb = 0.0, SRb = 0.0, SMb = 0.0, vnLb = 0.0, vnRb = 0.0;
    if (star) {
        const double S = left ? SL : SR;
        const double sv = S - vnW;
        const double dsm = S - SM;
        const bool dok = fabs(dsm) > kFloor;
        const double den1 = dok ? dsm : kFloor;
        const double fac = sv / den1;
        const double rs = W.r * fac;
        const double d = SM - vnW;
        const double us = W.u + d * nx, vs = W.v + d * ny;
        const double qv = W.r * sv;
        const bool qok = fabs(qv) > kFloor;
        const double q = qok ? qv : kFloor;
        const double Y = SM + W.p / q;
        const double Es = U[3] / W.r + d * Y;
        const double X[4] = {rs, rs * us, rs * vs, rs * Es};
        double Sb = 0.0, Xb[4];
        for (int k = 0; k < 4; ++k) {
            Sb += ob[k] * (X[k] - U[k]);
            Xb[k] = ob[k] * S;
            Ub[k] -= ob[k] * S;
        }
        const double rsb = Xb[0] + Xb[1] * us + Xb[2] * vs + Xb[3] * Es;
        const double usb = Xb[1] * rs, vsb = Xb[2] * rs, Esb = Xb[3] * rs;
        // Es = U3 / r + d Y,  Y = SM + p / q,  q = af(r (S - vn))
        Ub[3] += Esb / W.r;
        Wb[0] -= Esb * U[3] / (W.r * W.r);
        double db = Esb * Y + usb * nx + vsb * ny;
        const double Yb = Esb * d;
        SMb += Yb;
        Wb[3] += Yb / q;
        const double qb = -Yb * W.p / (q * q);
        double svb = 0.0;
        if (qok) {
            Wb[0] += qb * sv;
            svb += qb * W.r;
        }
        Wb[1] += usb;
        Wb[2] += vsb;
        // rs = r fac,  fac = (S - vn) / af(S - SM)
        Wb[0] += rsb * fac;
        const double facb = rsb * W.r;
        svb += facb / den1;
        const double den1b = -facb * sv / (den1 * den1);
        if (dok) {
            Sb += den1b;
            SMb -= den1b;
        }
        // d = SM - vn,  sv = S - vn
        SMb += db;
        Sb += svb;
        const double vnb = -db - svb;
        if (left) {
            SLb += Sb;
            vnLb += vnb;
        } else {
            SRb += Sb;
            vnRb += vnb;
        }
        // SM = num / af(mR - mL)
        const double numb = SMb / den;
        const double denb = -SMb * num / (den * den);
        double mLb = 0.0, mRb = 0.0;
        if (dmok) {
            mRb += denb;
            mLb -= denb;
        }
        Rb[0] += numb * vnR * (SR - vnR);
        vnRb += numb * R.r * (SR - 2.0 * vnR);
        SRb += numb * R.r * vnR;
        Lb[0] -= numb * vnL * (SL - vnL);
        vnLb -= numb * L.r * (SL - 2.0 * vnL);
        SLb -= numb * L.r * vnL;
        Lb[3] += numb;
        Rb[3] -= numb;
        // mL = rL (SL - vnL), mR = rR (SR - vnR)
        Lb[0] += mLb * (SL - vnL);
        SLb += mLb * L.r;
        vnLb -= mLb * L.r;
        Rb[0] += mRb * (SR - vnR);
        SRb += mRb * R.r;
        vnRb -= mRb * R.r;
    }
    // wave speeds, the low-Mach blend first
    double SL0b = SLb, SR0b = SRb, vntb = 0.0, aLb = 0.0, aRb = 0.0;
    if (lm) {
        SL0b = SLb * phi;
        SR0b = SRb * phi;
        const double lob = SLb * (1.0 - phi), hib = SRb * (1.0 - phi);
        const double phib = -SLb * (lo - SL0) + SRb * (SR0 - hi);
        if (loA) vnLb += lob;
        else vntb += lob;
        if (hiA) vnRb += hib;
        else vntb += hib;
        if (!phiA) {
            const double machb = phib / low_mach;
            if (machA) {
                vnLb += machb * (vnL < 0.0 ? -1.0 : 1.0) / aL;
                aLb -= machb * xL / aL;
            } else {
                vnRb += machb * (vnR < 0.0 ? -1.0 : 1.0) / aR;
                aRb -= machb * xR / aR;
            }
        }
    }
    double atb = 0.0;
    if (slA) {
        vnLb += SL0b;
        aLb -= SL0b;
    } else {
        vntb += SL0b;
        atb -= SL0b;
    }
    if (srA) {
        vnRb += SR0b;
        aRb += SR0b;
    } else {
        vntb += SR0b;
        atb += SR0b;
    }
    // at = sqrt(max((g - 1)(H - (vx^2 + vy^2) / 2), floor)), vnt = vx nx + vy ny
    double vxb = vntb * nx, vyb = vntb * ny, Hb = 0.0;
    if (cok) {
        const double cqb = atb * 0.5 / at;
        Hb += (g - 1.0) * cqb;
        vxb -= (g - 1.0) * cqb * vx;
        vyb -= (g - 1.0) * cqb * vy;
    }
    // Roe means, as in roe_vjp
    const double nvxb = vxb * iden, nvyb = vyb * iden, nHb = Hb * iden;
    const double idenb = vxb * nvx + vyb * nvy + Hb * nH;
    const double sumb = -idenb * iden * iden;
    const double sLb = sumb + nvxb * L.u + nvyb * L.v + nHb * L.H;
    const double sRb = sumb + nvxb * R.u + nvyb * R.v + nHb * R.H;
    Lb[0] += sLb * 0.5 / sL;
    Rb[0] += sRb * 0.5 / sR;
    Lb[1] += nvxb * sL;
    Lb[2] += nvyb * sL;
    Lb[4] += nHb * sL;
    Rb[1] += nvxb * sR;
    Rb[2] += nvyb * sR;
    Rb[4] += nHb * sR;
    // a = sqrt(max(g p ir, floor)), vn = u nx + v ny
    if (zLok) {
        const double zb = aLb * 0.5 / aL;
        Lb[3] += zb * g * L.ir;
        Lb[5] += zb * g * L.p;
    }
    if (zRok) {
        const double zb = aRb * 0.5 / aR;
        Rb[3] += zb * g * R.ir;
        Rb[5] += zb * g * R.p;
    }
    Lb[1] += vnLb * nx;
    Lb[2] += vnLb * ny;
    Rb[1] += vnRb * nx;
    Rb[2] += vnRb * ny;
    // the own side's primitives back to its conserved state
    const double* Ob = own_left ? Lb : Rb;
    const Prim<double>& O = own_left ? L : R;
    primitives_vjp(own_left ? UL : UR, g, Ob[0] - Ob[5] * O.ir * O.ir, Ob[1], Ob[2], Ob[3], Ob[4],
                   out);
    if (own_left == left)
        for (int k = 0; k < 4; ++k) out[k] += Ub[k];
}

// FLUX: 0 roe, 1 hllc, 2 ausmup (src/backends/numba_backend.py:59-67)
template <int FLUX, class T, class N_>
__device__ __forceinline__ T interface_flux(const nz_problem& P, const T UL[4], const T UR[4],
                                            const N_& nx, const N_& ny, T out[4]) {
    if constexpr (FLUX == 1) return flux_hllc(UL, UR, nx, ny, P.gamma, P.low_mach, out);
    else if constexpr (FLUX == 2) return flux_ausm(UL, UR, nx, ny, P.gamma, out);
    else return flux_roe(UL, UR, nx, ny, P.gamma, P.efix, out);
}

// tag: 1 inflow, 2 outflow, 3 wall, 4 axis (src/geometry/mesh.py:34-40)
template <class T>
__device__ __forceinline__ T boundary_flux(const nz_problem& P, int tag, const T U[4], const T& nx,
                                           const T& ny, const T& pb, T out[4]) {
    if (tag == 1) return flux_inflow(U, nx, ny, P, out);
    if (tag == 2) return flux_outflow(U, nx, ny, P, pb, out);
    return flux_wall(U, nx, ny, P.gamma, out);
}

}  // namespace nz
