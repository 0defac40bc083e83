/*
 * CPU oracle for the DG residual / SSP-RK3 hot path.  TEST INFRASTRUCTURE ONLY:
 * it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline leg, always as the checker or the timed CPU reference, never as
 * part of the product path.
 *
 * A loop-for-loop restatement of the reference's compiled CPU kernels
 * (src/backends/_numba_kernels.py), with the formula forms of its NumPy
 * oracle path where the two differ (src/physics/fluxes.py,
 * src/physics/boundary.py, src/spatial/limiter.py).  Arrays use the
 * reference's dense element-batched layout (src/spatial/operators.py:138-190):
 *   U, out          (E, N, 4)           fw      (F, nqf, 4)
 *   phi_vol         (N, nqv)            phi_face (2, nface, N, nqf)
 *   grad_x/grad_y   (E, N, nqv)         face_basis_m (E, nface, N, nqf)
 *   inv_mass        (E, N, N)           edge_normal (F, nqf, 2), edge_jac (F, nqf)
 * Index arrays are int64 as in the reference.
 *
 * Parity of every function here is pinned by tests/test_oracle.py against
 * golden vectors written from the reference's NumPy FP64 path
 * (tests/golden/make_goldens.py).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#define FLOOR 1e-10
#define TAG_INFLOW 1
#define TAG_OUTFLOW 2

typedef long long i64;

typedef struct {
    double gamma, efix, low_mach;
    int flux_id; /* 0 roe, 1 hllc, 2 ausmup (src/backends/numba_backend.py:59) */
    double rho_t_bf, p_t_bf, band_bf;        /* backflow branch; rho_t_bf = 0 disables */
    double at2, at, rho_t, ca, sa, p_back;   /* inflow/outflow constants */
} or_flow;

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* primitives: src/physics/thermo.py:46-53 */
static inline void prim(const double *U, double g, double *rho, double *vx, double *vy,
                        double *p, double *H) {
    double r = dmax(U[0], FLOOR);
    double u = U[1] / r, v = U[2] / r;
    double pp = dmax((g - 1.0) * (U[3] - 0.5 * r * (u * u + v * v)), FLOOR);
    *rho = r; *vx = u; *vy = v; *p = pp; *H = (U[3] + pp) / r;
}

/* normal_flux: thermo.py:80-86 */
static inline void nflux(const double *U, double nx, double ny, double g, double *f) {
    double r, u, v, p, H;
    prim(U, g, &r, &u, &v, &p, &H);
    double vn = u * nx + v * ny;
    f[0] = r * vn;
    f[1] = r * u * vn + p * nx;
    f[2] = r * v * vn + p * ny;
    f[3] = r * H * vn;
}

/* roe_flux: fluxes.py:38-103 */
static double roe(const double *UL, const double *UR, double nx, double ny, double g,
                  double efix, double *out) {
    double rL, uL, vL, pL, HL, rR, uR, vR, pR, HR;
    prim(UL, g, &rL, &uL, &vL, &pL, &HL);
    prim(UR, g, &rR, &uR, &vR, &pR, &HR);
    double FL[4], FR[4];
    nflux(UL, nx, ny, g, FL);
    nflux(UR, nx, ny, g, FR);
    double sL = sqrt(rL), sR = sqrt(rR), den = sL + sR;
    double vx = (sL * uL + sR * uR) / den;
    double vy = (sL * vL + sR * vR) / den;
    double H = (sL * HL + sR * HR) / den;
    double q2 = vx * vx + vy * vy;
    double c = sqrt(dmax((g - 1.0) * (H - 0.5 * q2), FLOOR));
    double vn = vx * nx + vy * ny;
    double l1 = fabs(vn + c), l2 = fabs(vn - c), l3 = fabs(vn);
    double smax = dmax(l1, l2);
    double eps = efix * c;
    if (l1 < eps) l1 = (l1 * l1 + eps * eps) / (2.0 * eps);
    if (l2 < eps) l2 = (l2 * l2 + eps * eps) / (2.0 * eps);
    if (l3 < eps) l3 = (l3 * l3 + eps * eps) / (2.0 * eps);
    double d0 = UR[0] - UL[0], d1 = UR[1] - UL[1], d2 = UR[2] - UL[2], d3 = UR[3] - UL[3];
    double G1 = (g - 1.0) * (0.5 * q2 * d0 - vx * d1 - vy * d2 + d3);
    double G2 = -vn * d0 + d1 * nx + d2 * ny;
    double s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2);
    double C1 = (G1 / (c * c)) * (s1 - l3) + (G2 / c) * s2;
    double C2 = (G1 / c) * s2 + (s1 - l3) * G2;
    out[0] = 0.5 * (FL[0] + FR[0]) - 0.5 * (l3 * d0 + C1);
    out[1] = 0.5 * (FL[1] + FR[1]) - 0.5 * (l3 * d1 + C1 * vx + C2 * nx);
    out[2] = 0.5 * (FL[2] + FR[2]) - 0.5 * (l3 * d2 + C1 * vy + C2 * ny);
    out[3] = 0.5 * (FL[3] + FR[3]) - 0.5 * (l3 * d3 + C1 * H + C2 * vn);
    return smax;
}

static inline double guard(double x) { return fabs(x) > FLOOR ? x : FLOOR; }

/* hllc_flux: fluxes.py:106-255 */
static double hllc(const double *UL, const double *UR, double nx, double ny, double g,
                   double low_mach, double *out) {
    double rL, uL, vL, pL, HL, rR, uR, vR, pR, HR;
    prim(UL, g, &rL, &uL, &vL, &pL, &HL);
    prim(UR, g, &rR, &uR, &vR, &pR, &HR);
    double vnL = uL * nx + vL * ny, vnR = uR * nx + vR * ny;
    double aL = sqrt(dmax(g * pL / rL, FLOOR)), aR = sqrt(dmax(g * pR / rR, FLOOR));
    double sL = sqrt(rL), sR = sqrt(rR), den = sL + sR;
    double vx = (sL * uL + sR * uR) / den;
    double vy = (sL * vL + sR * vR) / den;
    double H = (sL * HL + sR * HR) / den;
    double vnt = vx * nx + vy * ny;
    double at = sqrt(dmax((g - 1.0) * (H - 0.5 * (vx * vx + vy * vy)), FLOOR));
    double SL = dmin(vnL - aL, vnt - at);
    double SR = dmax(vnR + aR, vnt + at);
    if (low_mach > 0.0) {
        double mach = dmax(fabs(vnL) / aL, fabs(vnR) / aR);
        double phi = dmin(1.0, mach / low_mach);
        double lo = dmin(vnL, vnt), hi = dmax(vnR, vnt);
        SL = lo - phi * (lo - SL);
        SR = hi + phi * (SR - hi);
    }
    double mL = rL * (SL - vnL), mR = rR * (SR - vnR);
    double SM = (rR * vnR * (SR - vnR) - rL * vnL * (SL - vnL) + pL - pR) / guard(mR - mL);
    double FL[4], FR[4];
    nflux(UL, nx, ny, g, FL);
    nflux(UR, nx, ny, g, FR);
    double ms = dmax(fabs(SL), fabs(SR));
    if (SL >= 0.0) { memcpy(out, FL, sizeof FL); return ms; }
    if (SR <= 0.0) { memcpy(out, FR, sizeof FR); return ms; }
    const double *U = SM >= 0.0 ? UL : UR;
    const double *F = SM >= 0.0 ? FL : FR;
    double r = SM >= 0.0 ? rL : rR, u = SM >= 0.0 ? uL : uR, v = SM >= 0.0 ? vL : vR;
    double p = SM >= 0.0 ? pL : pR, vn = SM >= 0.0 ? vnL : vnR, S = SM >= 0.0 ? SL : SR;
    double fac = (S - vn) / guard(S - SM);
    double rs = r * fac;
    double us = u + (SM - vn) * nx, vs = v + (SM - vn) * ny;
    double E = U[3] / r;
    double Es = E + (SM - vn) * (SM + p / guard(r * (S - vn)));
    double Us[4] = {rs, rs * us, rs * vs, rs * Es};
    for (int s = 0; s < 4; ++s) out[s] = F[s] + S * (Us[s] - U[s]);
    return ms;
}

/* AUSM+-up split polynomials: fluxes.py:275-318 (NumPy forms) */
static inline double m4p(double m) {
    double m2p = 0.25 * ((m + 1.0) * (m + 1.0)), m2m = -0.25 * ((m - 1.0) * (m - 1.0));
    return fabs(m) >= 1.0 ? 0.5 * (m + fabs(m)) : m2p * (1.0 - 16.0 * 0.125 * m2m);
}
static inline double m4m(double m) {
    double m2p = 0.25 * ((m + 1.0) * (m + 1.0)), m2m = -0.25 * ((m - 1.0) * (m - 1.0));
    return fabs(m) >= 1.0 ? 0.5 * (m - fabs(m)) : m2m * (1.0 + 16.0 * 0.125 * m2p);
}
static inline double sgn(double m) { return m > 0.0 ? 1.0 : (m < 0.0 ? -1.0 : 0.0); }
static inline double p5p(double m) {
    double m2p = 0.25 * ((m + 1.0) * (m + 1.0)), m2m = -0.25 * ((m - 1.0) * (m - 1.0));
    return fabs(m) >= 1.0 ? 0.5 * (1.0 + sgn(m)) : m2p * ((2.0 - m) - 16.0 * 0.1875 * m * m2m);
}
static inline double p5m(double m) {
    double m2p = 0.25 * ((m + 1.0) * (m + 1.0)), m2m = -0.25 * ((m - 1.0) * (m - 1.0));
    return fabs(m) >= 1.0 ? 0.5 * (1.0 - sgn(m)) : m2m * ((-2.0 - m) + 16.0 * 0.1875 * m * m2p);
}

/* ausm_up_flux: fluxes.py:321-445 */
static double ausm(const double *UL, const double *UR, double nx, double ny, double g,
                   double *out) {
    double rL, uL, vL, pL, HL, rR, uR, vR, pR, HR;
    prim(UL, g, &rL, &uL, &vL, &pL, &HL);
    prim(UR, g, &rR, &uR, &vR, &pR, &HR);
    double vnL = uL * nx + vL * ny, vnR = uR * nx + vR * ny;
    double aL = sqrt(g * pL / rL), aR = sqrt(g * pR / rR);
    double ah = 0.5 * (aL + aR), rh = 0.5 * (rL + rR);
    double mL = vnL / ah, mR = vnR / ah;
    double mb2 = 0.5 * (mL * mL + mR * mR);
    double uh = ah * (m4p(mL) + m4m(mR)) + 0.4 * dmax(1.0 - mb2, 0.0) * (pL - pR) / (rh * ah);
    double pp = p5p(mL), pm = p5m(mR);
    double ph = pp * pL + pm * pR + 0.2 * pp * pm * rh * ah * (vnL - vnR);
    double PL[4] = {rL, rL * uL, rL * vL, UL[3] + pL};
    double PR[4] = {rR, rR * uR, rR * vR, UR[3] + pR};
    double au = fabs(uh);
    for (int s = 0; s < 4; ++s) out[s] = 0.5 * uh * (PL[s] + PR[s]) - 0.5 * au * (PR[s] - PL[s]);
    out[1] = out[1] + ph * (nx + 0.0);
    out[2] = out[2] + ph * (ny + 0.0);
    out[0] = out[0] + ph * 0.0;
    out[3] = out[3] + ph * 0.0;
    return dmax(fabs(vnL) + aL, fabs(vnR) + aR);
}

/* wall_flux: boundary.py:20-38 */
static double wall(const double *U, double nx, double ny, double g, double *out) {
    double r = dmax(U[0], FLOOR), u = U[1] / r, v = U[2] / r;
    double vn = u * nx + v * ny;
    double tx = u - vn * nx, ty = v - vn * ny;
    double pb = dmax((g - 1.0) * (U[3] - 0.5 * r * (tx * tx + ty * ty)), FLOOR);
    out[0] = 0.0; out[1] = pb * nx; out[2] = pb * ny; out[3] = 0.0;
    return sqrt(g * pb / r);
}

/* inflow_flux: boundary.py:41-110 */
static double inflow(const double *U, double nx, double ny, const or_flow *fl, double *out) {
    double g = fl->gamma, r, u, v, p, H;
    prim(U, g, &r, &u, &v, &p, &H);
    double a = sqrt(g * p / r);
    double vn = u * nx + v * ny;
    double Jp = vn + 2.0 * a / (g - 1.0);
    double beta = (Jp / fl->at) * (Jp / fl->at);
    double nd = nx * fl->ca + ny * fl->sa;
    double aa = 0.5 * (g - 1.0) * beta - nd * nd;
    double bb = -4.0 * nd / (g - 1.0);
    double t2 = 2.0 / (g - 1.0);
    double cc = beta - t2 * t2;
    if (fabs(aa) < 1e-12) aa = aa >= 0.0 ? 1e-12 : -1e-12;
    double disc = dmax(bb * bb - 4.0 * aa * cc, 0.0);
    double sq = sqrt(disc);
    double r1 = (-bb + sq) / (2.0 * aa), r2 = (-bb - sq) / (2.0 * aa);
    double c1 = r1 >= 0.0 ? r1 : 1e30, c2 = r2 >= 0.0 ? r2 : 1e30;
    double M = dmin(c1, c2);
    if (M > 0.5e30) M = 0.0;
    M = dmin(dmax(M, 0.0), 1.0);
    double fac = 1.0 + 0.5 * (g - 1.0) * M * M;
    double ab = sqrt(fl->at2 / fac);
    double qb = M * ab, vxb = qb * fl->ca, vyb = qb * fl->sa;
    double rhob = fl->rho_t * pow(fac, -1.0 / (g - 1.0));
    double pb = rhob * ab * ab / g;
    double Hb = fl->at2 / (g - 1.0);
    double vnb = vxb * nx + vyb * ny;
    out[0] = rhob * vnb;
    out[1] = rhob * vxb * vnb + pb * nx;
    out[2] = rhob * vyb * vnb + pb * ny;
    out[3] = rhob * Hb * vnb;
    return fabs(vnb) + ab;
}

/* outflow_flux: boundary.py:113-214 (NumPy form, incl. its max-speed definition) */
static double outflow(const double *U, double nx, double ny, const or_flow *fl, double *out) {
    double g = fl->gamma, r, u, v, p, H;
    prim(U, g, &r, &u, &v, &p, &H);
    double a = sqrt(g * p / r);
    double vn = u * nx + v * ny;
    double Ub[4];
    if (vn / a >= 1.0) {
        memcpy(Ub, U, sizeof Ub);
    } else {
        double pb = fl->p_back;
        double rhob = r * pow(pb / p, 1.0 / g);
        double ab = sqrt(g * pb / rhob);
        double vnb = vn + 2.0 / (g - 1.0) * (a - ab);
        double vxb = (u - vn * nx) + vnb * nx, vyb = (v - vn * ny) + vnb * ny;
        double Eb = pb / ((g - 1.0) * rhob) + 0.5 * (vxb * vxb + vyb * vyb);
        Ub[0] = rhob; Ub[1] = rhob * vxb; Ub[2] = rhob * vyb; Ub[3] = rhob * Eb;
        if (fl->rho_t_bf > 0.0) {
            double rf = fl->rho_t_bf * pow(pb / fl->p_t_bf, 1.0 / g);
            double af = sqrt(g * pb / rf);
            double vnf = vn + 2.0 / (g - 1.0) * (a - af);
            double vxf = vnf * nx, vyf = vnf * ny;
            double Ef = pb / ((g - 1.0) * rf) + 0.5 * (vxf * vxf + vyf * vyf);
            double Uf[4] = {rf, rf * vxf, rf * vyf, rf * Ef};
            double t = dmin(dmax((vn / a) / dmax(fl->band_bf, 1e-12), 0.0), 1.0);
            for (int s = 0; s < 4; ++s) Ub[s] = t * Ub[s] + (1.0 - t) * Uf[s];
        }
    }
    nflux(Ub, nx, ny, g, out);
    double ro, uo, vo, po, Ho;
    prim(Ub, g, &ro, &uo, &vo, &po, &Ho);
    return fabs(uo * nx + vo * ny) + sqrt(g * po / ro);
}

/* edge_pass: _numba_kernels.py:484-583 */
void or_edge_pass(i64 n_int, i64 n_bnd, int nbf, int nqf, int nface, const double *U,
                  const double *phi_face, const i64 *iedge_elem, const i64 *iedge_face,
                  const i64 *bedge_elem, const i64 *bedge_face, const i64 *bedge_tag,
                  const double *edge_normal, const double *edge_jac, const double *w_face,
                  const or_flow *fl, double *fw, double *smax) {
    const double *phi0 = phi_face, *phi1 = phi_face + (size_t)nface * nbf * nqf;
#pragma omp parallel for schedule(static)
    for (i64 k = 0; k < n_int + n_bnd; ++k) {
        double UL[4], UR[4], f[4], best = 0.0;
        for (int q = 0; q < nqf; ++q) {
            double nx = edge_normal[(k * nqf + q) * 2], ny = edge_normal[(k * nqf + q) * 2 + 1];
            double sp;
            if (k < n_int) {
                i64 le = iedge_elem[2 * k], re = iedge_elem[2 * k + 1];
                i64 lf = iedge_face[2 * k], rf = iedge_face[2 * k + 1];
                for (int s = 0; s < 4; ++s) UL[s] = UR[s] = 0.0;
                for (int i = 0; i < nbf; ++i) {
                    double bl = phi0[(lf * nbf + i) * nqf + q], br = phi1[(rf * nbf + i) * nqf + q];
                    for (int s = 0; s < 4; ++s) {
                        UL[s] += bl * U[(le * nbf + i) * 4 + s];
                        UR[s] += br * U[(re * nbf + i) * 4 + s];
                    }
                }
                if (fl->flux_id == 1) sp = hllc(UL, UR, nx, ny, fl->gamma, fl->low_mach, f);
                else if (fl->flux_id == 2) sp = ausm(UL, UR, nx, ny, fl->gamma, f);
                else sp = roe(UL, UR, nx, ny, fl->gamma, fl->efix, f);
            } else {
                i64 b = k - n_int, be = bedge_elem[b], bf = bedge_face[b];
                for (int s = 0; s < 4; ++s) UL[s] = 0.0;
                for (int i = 0; i < nbf; ++i) {
                    double bl = phi0[(bf * nbf + i) * nqf + q];
                    for (int s = 0; s < 4; ++s) UL[s] += bl * U[(be * nbf + i) * 4 + s];
                }
                i64 tag = bedge_tag[b];
                if (tag == TAG_INFLOW) sp = inflow(UL, nx, ny, fl, f);
                else if (tag == TAG_OUTFLOW) sp = outflow(UL, nx, ny, fl, f);
                else sp = wall(UL, nx, ny, fl->gamma, f);
            }
            if (sp > best) best = sp;
            double sc = edge_jac[k * nqf + q] * w_face[q];
            for (int s = 0; s < 4; ++s) fw[(k * nqf + q) * 4 + s] = f[s] * sc;
        }
        smax[k] = best;
    }
}

static inline void euler_fg(const double *Uq, double g, double *F, double *G) {
    double r, u, v, p, H;
    prim(Uq, g, &r, &u, &v, &p, &H);
    F[0] = r * u; F[1] = r * u * u + p; F[2] = r * u * v; F[3] = r * u * H;
    G[0] = r * v; G[1] = r * u * v; G[2] = r * v * v + p; G[3] = r * v * H;
}

/* element_pass (apply_mass 0/1) and element_rate (folded, fold=1):
 * _numba_kernels.py:586-759.  With fold=1, gx/gy are grad_*_m and fb is
 * face_basis_m (E, nface, N, nqf); otherwise fb is phi_face and face_side
 * selects the side. */
void or_element_pass(i64 nelem, int nbf, int nqv, int nqf, int nface, const double *U,
                     const double *fw, const double *smax, const double *phi_vol,
                     const double *gx, const double *gy, const double *fb, int fold,
                     const i64 *face_edge, const i64 *face_side, const double *face_sign,
                     const double *edge_length, double gamma, const double *inv_mass,
                     int apply_mass, double *out, double *wave) {
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < nelem; ++e) {
        double Re[64 * 4];
        double Uq[4], F[4], G[4];
        for (int i = 0; i < nbf * 4; ++i) Re[i] = 0.0;
        for (int q = 0; q < nqv; ++q) {
            for (int s = 0; s < 4; ++s) Uq[s] = 0.0;
            for (int i = 0; i < nbf; ++i) {
                double b = phi_vol[i * nqv + q];
                for (int s = 0; s < 4; ++s) Uq[s] += b * U[(e * nbf + i) * 4 + s];
            }
            euler_fg(Uq, gamma, F, G);
            for (int i = 0; i < nbf; ++i) {
                double a = gx[(e * nbf + i) * nqv + q], c = gy[(e * nbf + i) * nqv + q];
                for (int s = 0; s < 4; ++s) Re[i * 4 + s] -= a * F[s] + c * G[s];
            }
        }
        double acc = 0.0;
        for (int f = 0; f < nface; ++f) {
            i64 k = face_edge[e * nface + f];
            double sg = face_sign[e * nface + f];
            for (int q = 0; q < nqf; ++q)
                for (int i = 0; i < nbf; ++i) {
                    double b;
                    if (fold) b = sg * fb[((e * nface + f) * nbf + i) * nqf + q];
                    else {
                        i64 sd = face_side[e * nface + f];
                        b = sg * fb[((sd * nface + f) * nbf + i) * nqf + q];
                    }
                    for (int s = 0; s < 4; ++s) Re[i * 4 + s] += b * fw[(k * nqf + q) * 4 + s];
                }
            acc += smax[k] * edge_length[k];
        }
        wave[e] = acc;
        if (fold) {
            for (int i = 0; i < nbf * 4; ++i) out[e * nbf * 4 + i] = -Re[i];
        } else if (apply_mass) {
            for (int i = 0; i < nbf; ++i)
                for (int s = 0; s < 4; ++s) {
                    double t = 0.0;
                    for (int j = 0; j < nbf; ++j) t += inv_mass[(e * nbf + i) * nbf + j] * Re[j * 4 + s];
                    out[(e * nbf + i) * 4 + s] = -t;
                }
        } else {
            for (int i = 0; i < nbf * 4; ++i) out[e * nbf * 4 + i] = Re[i];
        }
    }
}

/* apply_inverse_mass: _numba_kernels.py:762-777 */
void or_inverse_mass(i64 nelem, int nbf, const double *inv_mass, const double *R, double *out) {
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < nelem; ++e)
        for (int i = 0; i < nbf; ++i)
            for (int s = 0; s < 4; ++s) {
                double t = 0.0;
                for (int j = 0; j < nbf; ++j) t += inv_mass[(e * nbf + i) * nbf + j] * R[(e * nbf + j) * 4 + s];
                out[(e * nbf + i) * 4 + s] = t;
            }
}

/* rms_weighted: _numba_kernels.py:789-814 */
double or_rms_weighted(i64 nelem, int nbf, const double *A, const double *w) {
    double total = 0.0;
#pragma omp parallel for reduction(+ : total) schedule(static)
    for (i64 e = 0; e < nelem; ++e) {
        double acc = 0.0;
        for (int i = 0; i < nbf; ++i)
            for (int s = 0; s < 4; ++s) {
                double v = A[(e * nbf + i) * 4 + s] * w[s];
                acc += v * v;
            }
        total += acc;
    }
    return sqrt(total / ((double)nelem * nbf * 4));
}

/* local_dt: _numba_kernels.py:817-824 and step.py:171-183 */
void or_local_dt(i64 nelem, const double *wave, const double *area, double factor, double *dt) {
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < nelem; ++e) dt[e] = factor * 2.0 * area[e] / dmax(wave[e], FLOOR);
}

/* stage / combine_ssp: _numba_kernels.py:827-850 */
void or_stage(i64 nelem, int nbf, const double *U, const double *F, double coef, const double *dt,
              double *out) {
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < nelem; ++e) {
        double c = coef * dt[e];
        for (int i = 0; i < nbf * 4; ++i) out[e * nbf * 4 + i] = U[e * nbf * 4 + i] + c * F[e * nbf * 4 + i];
    }
}

void or_combine_ssp(i64 nelem, int nbf, const double *U, const double *V, const double *F,
                    const double *dt, double a, double b, double *out) {
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < nelem; ++e) {
        double d = dt[e];
        for (int i = 0; i < nbf * 4; ++i)
            out[e * nbf * 4 + i] = a * U[e * nbf * 4 + i] + b * (V[e * nbf * 4 + i] + d * F[e * nbf * 4 + i]);
    }
}

/* Minimum density and pressure over the probe set at scaling theta:
 * _numba_kernels.py:858-924 / limiter.py:140-148 */
static void probe_minima(const double *Ue, const double *ub, double theta, int nbf, int nqv,
                         int nqf, int nface, const double *phi_vol, const double *phi_face,
                         const i64 *fside, double g, double *rmin, double *pmin) {
    double rm = 1e300, pm = 1e300;
    int nprobe = nqv + nface * nqf;
    for (int k = 0; k < nprobe; ++k) {
        double V[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < nbf; ++i) {
            double b;
            if (k < nqv) b = phi_vol[i * nqv + k];
            else {
                int f = (k - nqv) / nqf, q = (k - nqv) % nqf;
                b = phi_face[((fside[f] * nface + f) * nbf + i) * nqf + q];
            }
            for (int s = 0; s < 4; ++s) V[s] += b * Ue[i * 4 + s];
        }
        for (int s = 0; s < 4; ++s) V[s] = ub[s] + theta * (V[s] - ub[s]);
        if (V[0] < rm) rm = V[0];
        double r = V[0] > 1e-300 ? V[0] : 1e-300;
        double pp = (g - 1.0) * (V[3] - 0.5 * (V[1] * V[1] + V[2] * V[2]) / r);
        if (pp < pm) pm = pp;
    }
    *rmin = rm; *pmin = pm;
}

/* positivity_limit, in place: _numba_kernels.py:927-1093, with the flag
 * definitions of limiter.py:125-129 (bit 0, mean repaired) and :222-224
 * (bit 1, probe bound violated).  The O(N) screen is skipped when
 * lebesgue <= 0.  counts = (n_scaled, n_repaired). */
void or_positivity_limit(i64 nelem, int nbf, int nqv, int nqf, int nface, double *U,
                         const double *phi_vol, const double *phi_face, const i64 *face_side,
                         const double *mean_weights, double g, double fraction, double rho_floor,
                         double p_floor, int steps, double lebesgue, unsigned char *flags,
                         i64 *counts) {
    i64 n_scaled = 0, n_repair = 0;
    double g1 = g - 1.0;
#pragma omp parallel for reduction(+ : n_scaled, n_repair) schedule(static)
    for (i64 e = 0; e < nelem; ++e) {
        double *Ue = U + e * nbf * 4;
        double ub[4] = {0, 0, 0, 0};
        unsigned char fl = 0;
        for (int i = 0; i < nbf; ++i)
            for (int s = 0; s < 4; ++s) ub[s] += mean_weights[e * nbf + i] * Ue[i * 4 + s];
        double rho_new = ub[0] > rho_floor ? ub[0] : rho_floor;
        double e_min = p_floor / g1 + 0.5 * (ub[1] * ub[1] + ub[2] * ub[2]) / rho_new;
        double rhoE_new = ub[3] > e_min ? ub[3] : e_min;
        double d0 = rho_new - ub[0], d3 = rhoE_new - ub[3];
        if (d0 != 0.0 || d3 != 0.0) {
            for (int i = 0; i < nbf; ++i) { Ue[i * 4] += d0; Ue[i * 4 + 3] += d3; }
            ub[0] = rho_new; ub[3] = rhoE_new;
            n_repair += 1;
            fl |= 1;
        }
        double p_bar = g1 * (ub[3] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2]) / ub[0]);
        double eps_rho = dmax(fraction * ub[0], 0.5 * rho_floor);
        double eps_p = dmax(fraction * p_bar, 0.5 * p_floor);
        int screened = 0;
        if (lebesgue > 0.0) {
            double dr = 0, dmx = 0, dmy = 0, den = 0;
            for (int i = 0; i < nbf; ++i) {
                dr = dmax(dr, fabs(Ue[i * 4] - ub[0]));
                dmx = dmax(dmx, fabs(Ue[i * 4 + 1] - ub[1]));
                dmy = dmax(dmy, fabs(Ue[i * 4 + 2] - ub[2]));
                den = dmax(den, fabs(Ue[i * 4 + 3] - ub[3]));
            }
            double rho_lo = ub[0] - lebesgue * dr;
            if (rho_lo >= eps_rho) {
                double m_hi = sqrt(ub[1] * ub[1] + ub[2] * ub[2]) + lebesgue * sqrt(dmx * dmx + dmy * dmy);
                double p_lo = g1 * (ub[3] - lebesgue * den - 0.5 * m_hi * m_hi / rho_lo);
                if (p_lo >= eps_p) screened = 1;
            }
        }
        if (!screened) {
            const i64 *fside = face_side + e * nface;
            double rmin, pmin;
            probe_minima(Ue, ub, 1.0, nbf, nqv, nqf, nface, phi_vol, phi_face, fside, g, &rmin, &pmin);
            if (!(rmin >= eps_rho && pmin >= eps_p)) {
                fl |= 2;
                double theta = 1.0;
                if (rmin < eps_rho) {
                    double gap = ub[0] - rmin;
                    theta = gap <= 0.0 ? 0.0 : dmin(dmax((ub[0] - eps_rho) / gap, 0.0), 1.0);
                }
                double rr, phi;
                probe_minima(Ue, ub, theta, nbf, nqv, nqf, nface, phi_vol, phi_face, fside, g, &rr, &phi);
                if (phi < eps_p) {
                    double lo = 0.0, hi = theta;
                    for (int it = 0; it < steps; ++it) {
                        double mid = 0.5 * (lo + hi), pm;
                        probe_minima(Ue, ub, mid, nbf, nqv, nqf, nface, phi_vol, phi_face, fside, g, &rr, &pm);
                        if (pm >= eps_p) lo = mid; else hi = mid;
                    }
                    theta = lo;
                }
                for (int i = 0; i < nbf; ++i)
                    for (int s = 0; s < 4; ++s) Ue[i * 4 + s] = ub[s] + theta * (Ue[i * 4 + s] - ub[s]);
                n_scaled += 1;
            }
        }
        if (flags) flags[e] = fl;
    }
    counts[0] = n_scaled;
    counts[1] = n_repair;
}

/* superbee(a, b) = maxmod(minmod(2a, b), minmod(a, 2b)): limiter.py:246-270 */
static inline double minmod(double a, double b) {
    return a * b <= 0.0 ? 0.0 : (fabs(a) < fabs(b) ? a : b);
}
static inline double maxmod(double a, double b) {
    return a * b <= 0.0 ? 0.0 : (fabs(a) > fabs(b) ? a : b);
}

/* superbee_limiter (without the positivity pass): limiter.py:286-405.
 * ubar (E,4) and trace means (E,nface,4) are computed here; the two global
 * reductions (total area, max |ubar_s|) are serial. */
void or_superbee(i64 nelem, int nbf, int nqf, int nface, double *U, const double *mean_weights,
                 const double *face_basis, const double *w_face, const i64 *face_neighbour,
                 const double *elem_area, double tvb, double *ubar_scratch) {
    const double tol = 1e-12;
    double *ub = ubar_scratch;
    for (i64 e = 0; e < nelem; ++e)
        for (int s = 0; s < 4; ++s) {
            double t = 0.0;
            for (int i = 0; i < nbf; ++i) t += mean_weights[e * nbf + i] * U[(e * nbf + i) * 4 + s];
            ub[e * 4 + s] = t;
        }
    double wsum = 0.0, area = 0.0, scale[4] = {0, 0, 0, 0};
    for (int q = 0; q < nqf; ++q) wsum += w_face[q];
    for (i64 e = 0; e < nelem; ++e) {
        area += elem_area[e];
        for (int s = 0; s < 4; ++s) scale[s] = dmax(scale[s], fabs(ub[e * 4 + s]));
    }
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < nelem; ++e) {
        double theta = 1.0;
        double share = elem_area[e] / area;
        for (int f = 0; f < nface; ++f) {
            i64 nb = face_neighbour[e * nface + f];
            for (int s = 0; s < 4; ++s) {
                double tm = 0.0;
                for (int i = 0; i < nbf; ++i)
                    for (int q = 0; q < nqf; ++q)
                        tm += face_basis[((e * nface + f) * nbf + i) * nqf + q] * w_face[q] * U[(e * nbf + i) * 4 + s];
                double dev = tm / wsum - ub[e * 4 + s];
                double jump = ub[nb * 4 + s] - ub[e * 4 + s];
                double allowed = maxmod(minmod(2.0 * jump, dev), minmod(jump, 2.0 * dev));
                double safe = fabs(dev) > tol ? dev : 1.0;
                double tf = dmin(dmax(allowed / safe, 0.0), 1.0);
                double thr = tvb * share * scale[s];
                int quiet = fabs(dev) <= tol || nb == e || fabs(dev) <= thr;
                if (quiet) tf = 1.0;
                if (tf < theta) theta = tf;
            }
        }
        for (int i = 0; i < nbf; ++i)
            for (int s = 0; s < 4; ++s) {
                double *u = &U[(e * nbf + i) * 4 + s];
                *u = ub[e * 4 + s] + theta * (*u - ub[e * 4 + s]);
            }
    }
}

/* OpenMP threads of the parallel loops above (timing of the CPU baseline at
 * one thread and at every core); returns the previous maximum. */
int or_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}
