// MI355X (gfx950) kernels for the DG residual, the fused SSP-RK3/Euler stage,
// the Zhang-Shu positivity limiter and the residual VJP, plus the C ABI
// declared in include/nozzle_hip.h.
//
// Work decomposition.
//
//   k_dg    the residual / rate / fused stage.  One wave (64 lanes) per
//           workgroup owns a tile of 16 elements, four lanes per element (one
//           per conserved variable).  Everything linear in U (traces,
//           volume-point interpolation, lift, projection, the stage update,
//           the limiter's probe sums) is separable by component, so each lane
//           keeps its element's column U[:, s] and its accumulator in
//           registers and reads reference tables with scalar loads; only the
//           pointwise physics (Riemann and Euler fluxes, one (point, element)
//           item per lane) goes through LDS.  Every interior flux is
//           evaluated by both elements of the edge with identical operands,
//           so both see the same bits and the assembly is a pure gather with
//           no flux table in HBM.  See the comment above k_dg.
//   k_vjp   the residual VJP and the adjoint RK4 stage.  One workgroup of 256
//           threads (four waves) owns a tile of TE elements (64, or 128 when
//           N <= 4), staged in LDS as [component][element] rows; each phase
//           maps the wave to a wave-uniform index (component, point or basis
//           function) and the lane to the element.  It serves the launch
//           that also returns geometry cotangents, and the state-cotangent
//           launches when the tile-shape autotuner picks four-wave tiles.
//   k_vjpw  the state-cotangent VJP and the adjoint RK4 stage on k_dg's
//           mapping (one wave per 16-element tile, four lanes per element).
//   others  small elementwise or per-element kernels (boundary pre-passes,
//           standalone limiter, superbee passes, inverse mass, time step,
//           norm finishing) on 256-thread workgroups.
//
// Affine triangles (Q = 1) fold M^-1 = M_ref^-1 / det into reference tables,
// so an element carries only (xi_x, xi_y, eta_x, eta_y, det).  Curved or
// quadrilateral elements carry per-point metrics and apply their own M^-1,
// stored as its upper triangle.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "nozzle_hip.h"
#include "physics.hip.h"

namespace nz {

constexpr int BS = 256;
constexpr int BJAC_DOUBLES = 20;  // per boundary face point of nz_problem.bjac

enum Mode { kRes = 0, kRate = 1, kStageFirst = 2, kStageNext = 3 };

template <int N_, int NQV_, int NQF_, int NF_, bool AFF_, bool STR_>
struct Cfg {
    static constexpr int N = N_, NQV = NQV_, NQF = NQF_, NF = NF_;
    static constexpr bool AFF = AFF_, STR = STR_;
    static constexpr int NS = N * 4;
    static constexpr int TE = N <= 4 ? 128 : 64;
    static constexpr int TP = TE + 1;          // padded LDS row (doubles)
    static constexpr int SLOTS = BS / TE;      // wave-uniform item slots per pass
    static constexpr int NIT = (N + SLOTS - 1) / SLOTS;  // (e,i) items per thread
    static constexpr int QR = SLOTS;           // volume points per round
    // LDS work area: face round | volume round | per-element R (non-affine rate)
    static constexpr int W_FACE = (2 * NQF * 4 + NQF) * TP;
    static constexpr int W_VOL = 2 * QR * 4 * TP;
    static constexpr int W_R = AFF ? 0 : NS * TP;
    static constexpr int WORK = W_FACE > W_VOL ? (W_FACE > W_R ? W_FACE : W_R)
                                               : (W_VOL > W_R ? W_VOL : W_R);
    static constexpr int PF = 2 * NF * N * NQF;  // phi_face (and lift) table size
    // adjoint work area: face round (neighbour U and lambda traces, cotangents)
    static constexpr int A_FACE = 3 * NQF * 4 * TP;
    static constexpr int AWORK = A_FACE > NS * TP ? (A_FACE > W_VOL ? A_FACE : W_VOL)
                                                 : (NS * TP > W_VOL ? NS * TP : W_VOL);
    // LDS bytes of one adjoint workgroup (four waves, one per SIMD) and the
    // waves per SIMD that the CU's 160 KB allows, at most three
    static constexpr int ALDS = 8 * (2 * NS * TP + AWORK + PF + 5 * TE) + 4 * NF * TE;
    static constexpr int AWPE = (160 * 1024) / ALDS < 1   ? 1
                                : (160 * 1024) / ALDS > 3 ? 3
                                                          : (160 * 1024) / ALDS;
    // polynomial order and the Lagrange nodes on one face (p + 1 of them)
    static constexpr int PO = NF == 3 ? (N == 1 ? 0 : N == 3 ? 1 : N == 6 ? 2 : 3)
                                      : (N == 1 ? 0 : N == 4 ? 1 : N == 9 ? 2 : 3);
    static constexpr int FN = PO + 1;
    // Face traces from the face's nodes alone.  The table entries of the other
    // nodes are round-off of an identically vanishing function (at most
    // 2.6e-15 for triangles and for quadrilaterals up to p = 2); p = 3
    // quadrilaterals carry up to 4.3e-14 there and keep the full sums, and
    // p = 0 has no node on a face.
    static constexpr bool SPT = PO >= 1 && !(NF == 4 && PO == 3);
    // Large face tables (quadrilaterals at P3, curved triangles at P3, Q2
    // quadrilaterals at P2): the traces over every node and the lifts loop
    // over the face points instead of unrolling them, so that only one
    // point's column of the table is held in scalar registers at a time (the
    // unrolled sums hoisted every entry and spilled hundreds of them).
    static constexpr bool ROLL_T = !SPT && N * NQF > 32;
    static constexpr bool ROLL_L = N * NQF > 32;
};

// Lagrange node j (0..p) of face f, in the reference's node numbering (an
// equispaced lattice (ix, iy), row by row in iy; elements.py:57-75).  Every
// other node's basis function vanishes identically on face f.  Triangles:
// face 0 is ix + iy = p, face 1 is ix = 0, face 2 is iy = 0; quadrilaterals:
// face 0 is iy = 0, face 1 is ix = p, face 2 is iy = p, face 3 is ix = 0.
__host__ __device__ constexpr int face_node(int nface, int p, int f, int j) {
    if (nface == 3) {
        const int ix = f == 0 ? p - j : (f == 1 ? 0 : j);
        const int iy = f == 2 ? 0 : j;
        return iy * (p + 1) - iy * (iy - 1) / 2 + ix;
    }
    const int ix = f == 0 ? j : (f == 1 ? p : (f == 2 ? j : 0));
    const int iy = f == 0 ? 0 : (f == 1 ? j : (f == 2 ? p : j));
    return iy * (p + 1) + ix;
}

// Sums sum_i tab[i * NQF + q] * col[i] at the face points of face F (its
// side-0 table `tab`): over the face's nodes alone when C::SPT, in node
// order j = 0..p, else over every basis function.  Every trace of the
// element kernels goes through here, so a trace formed by an element and the
// same trace formed by its neighbour round alike.
template <class C, int F, int NC, class OUT>
__device__ __forceinline__ void trace_face(const double* tab, const double (*const col[NC])[C::N],
                                           OUT out) {
    constexpr int N = C::N, NQF = C::NQF;
    if constexpr (C::ROLL_T) {
#pragma unroll 1
        for (int q = 0; q < NQF; ++q) {
            double a[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int c = 0; c < NC; ++c) a[c] += tab[i * NQF + q] * (*col[c])[i];
            out(q, a);
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < NQF; ++q) {
        double a[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = 0.0;
        if constexpr (C::SPT) {
#pragma unroll
            for (int j = 0; j < C::FN; ++j) {
                const int i = face_node(C::NF, C::PO, F, j);
#pragma unroll
                for (int c = 0; c < NC; ++c) a[c] += tab[i * NQF + q] * (*col[c])[i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int c = 0; c < NC; ++c) a[c] += tab[i * NQF + q] * (*col[c])[i];
        }
        out(q, a);
    }
}

// trace_face for a wave-uniform runtime face f (one branch per face)
template <class C, int NC, class OUT>
__device__ __forceinline__ void trace_face_rt(int f, const double* phi_face,
                                              const double (*const col[NC])[C::N], OUT out) {
    constexpr int S = C::N * C::NQF;
    if (f == 0) {
        trace_face<C, 0, NC>(phi_face, col, out);
    } else if (f == 1) {
        trace_face<C, 1, NC>(phi_face + S, col, out);
    } else if constexpr (C::NF == 3) {
        trace_face<C, 2, NC>(phi_face + 2 * S, col, out);
    } else {
        if (f == 2) trace_face<C, 2, NC>(phi_face + 2 * S, col, out);
        else trace_face<C, 3, NC>(phi_face + 3 * S, col, out);
    }
}

// The same sums for a neighbour's face: `cn` holds the neighbour's
// coefficients at the nodes of its face key (C::FN of them, node order) when
// C::SPT, else all C::N.  `tab` is the key's side-0 table (wave-uniform
// pointer: scalar loads, or a per-lane LDS copy).
template <class C, int NC, class TAB, class OUT>
__device__ __forceinline__ void trace_nbr(int key, TAB tab, const double (*const cn[NC])[C::N],
                                          OUT out) {
    constexpr int NQF = C::NQF;
    constexpr int NI = C::SPT ? C::FN : C::N;
    if constexpr (C::ROLL_T) {
#pragma unroll 1
        for (int q = 0; q < NQF; ++q) {
            double a[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] = 0.0;
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const double b = tab(j * NQF + q);
#pragma unroll
                for (int c = 0; c < NC; ++c) a[c] += b * (*cn[c])[j];
            }
            out(q, a);
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < NQF; ++q) {
        double a[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = 0.0;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int i = C::SPT ? face_node(C::NF, C::PO, key, j) : j;
            const double b = tab(i * NQF + q);
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] += b * (*cn[c])[j];
        }
        out(q, a);
    }
}

// node of slot j of a neighbour column: face nodes of face `key` when C::SPT
template <class C>
__device__ __forceinline__ int nbr_node(int key, int j) {
    return C::SPT ? face_node(C::NF, C::PO, key, j) : j;
}

// The face quadrature weight of point q (a per-lane index): the weights are
// read with wave-uniform scalar loads and picked by a select chain, so no
// lane waits on a vector load for them.
template <class C>
__device__ __forceinline__ double face_weight(const double* __restrict__ T, const nz_problem& P,
                                              int q) {
    double w = T[P.off.w_face];
#pragma unroll
    for (int k = 1; k < C::NQF; ++k) w = q == k ? T[P.off.w_face + k] : w;
    return w;
}

// whether basis function i enters the traces of face f
template <class C>
__device__ __forceinline__ bool on_face(int f, int i) {
    if (!C::SPT) return true;
    bool on = false;
#pragma unroll
    for (int j = 0; j < C::FN; ++j) on = on || face_node(C::NF, C::PO, f, j) == i;
    return on;
}

// acc[i] += sum_q tab[i * NQF + q] fw[q] over the nodes of face F (all nodes
// unless C::SPT): the transpose of the face trace, with the side-0 table
// acc[i] += sum_q tab[i * NQF + q] fw[q] over the nodes of face F (all nodes
// unless C::SPT): the transpose of the face trace, with the side-0 table
template <class C, int F>
__device__ __forceinline__ void lift_face(const double* tab, const double (&fw)[C::NQF],
                                          double (&acc)[C::N]) {
    constexpr int NQF = C::NQF, NI = C::SPT ? C::FN : C::N;
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        const int i = C::SPT ? face_node(C::NF, C::PO, F, j) : j;
#pragma unroll
        for (int q = 0; q < NQF; ++q) acc[i] += tab[i * NQF + q] * fw[q];
    }
}

template <class C>
__device__ __forceinline__ void lift_face_rt(int f, const double* phi_face,
                                             const double (&fw)[C::NQF], double (&acc)[C::N]) {
    constexpr int S = C::N * C::NQF;
    if (f == 0) {
        lift_face<C, 0>(phi_face, fw, acc);
    } else if (f == 1) {
        lift_face<C, 1>(phi_face + S, fw, acc);
    } else if constexpr (C::NF == 3) {
        lift_face<C, 2>(phi_face + 2 * S, fw, acc);
    } else {
        if (f == 2) lift_face<C, 2>(phi_face + 2 * S, fw, acc);
        else lift_face<C, 3>(phi_face + 3 * S, fw, acc);
    }
}

// The same lift over every node with the face points in a loop (large
// tables, see C::ROLL_T): fwq(q) gives the weighted flux at point q.  Each
// acc[i] receives its NQF terms in point order, as in lift_face.
template <class C, class FW>
__device__ __forceinline__ void lift_face_rolled(const double* tab, FW fwq, double (&acc)[C::N]) {
#pragma unroll 1
    for (int q = 0; q < C::NQF; ++q) {
        const double w = fwq(q);
#pragma unroll
        for (int i = 0; i < C::N; ++i) acc[i] += tab[i * C::NQF + q] * w;
    }
}

// ---------------------------------------------------------------- limiter
// Zhang-Shu positivity limiter on one element whose coefficients sit at
// u[(i*4+s)*stride] (src/backends/_numba_kernels.py:927-1093; flags as in
// src/spatial/limiter.py:125-129,222-224).  No FP contraction, so the probe
// sums and the bound tests round exactly as the CPU oracle's.
struct LimTables {
    const double* phi_vol;   // (N, nqv)
    const double* phi_face;  // (2, nface, N, nqf)
};

template <class C>
__device__ double probe_min(const double* u, int stride, const double ub[4], double theta,
                            const LimTables& tb, const int side[C::NF], double g, double* pmin) {
#pragma clang fp contract(off)
    double rm = 1e300, pm = 1e300;
#pragma unroll 1
    for (int k = 0; k < C::NQV + C::NF * C::NQF; ++k) {
        // re-read the coefficients every probe instead of pinning N*4 of
        // them in registers for the whole kernel
        asm volatile("" ::: "memory");
        double V[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < C::N; ++i) {
            double b;
            if (k < C::NQV) {
                b = tb.phi_vol[i * C::NQV + k];
            } else {
                const int f = (k - C::NQV) / C::NQF, q = (k - C::NQV) % C::NQF;
                b = tb.phi_face[((side[f] * C::NF + f) * C::N + i) * C::NQF + q];
            }
            for (int s = 0; s < 4; ++s) V[s] += b * u[(i * 4 + s) * stride];
        }
        for (int s = 0; s < 4; ++s) V[s] = ub[s] + theta * (V[s] - ub[s]);
        rm = V[0] < rm ? V[0] : rm;
        const double r = V[0] > 1e-300 ? V[0] : 1e-300;
        const double pp = (g - 1.0) * (V[3] - 0.5 * (V[1] * V[1] + V[2] * V[2]) / r);
        pm = pp < pm ? pp : pm;
    }
    *pmin = pm;
    return rm;
}

template <class C>
__device__ uint8_t limit_element(double* u, int stride, const double* mw, const LimTables& tb,
                                 const int side[C::NF], const nz_problem& P, int* scaled) {
#pragma clang fp contract(off)
    const double g = P.gamma, g1 = g - 1.0;
    uint8_t fl = 0;
    double ub[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < C::N; ++i)
        for (int s = 0; s < 4; ++s) ub[s] += mw[i] * u[(i * 4 + s) * stride];
    const double rho_new = ub[0] > P.rho_floor ? ub[0] : P.rho_floor;
    const double e_min = P.p_floor / g1 + 0.5 * (ub[1] * ub[1] + ub[2] * ub[2]) / rho_new;
    const double rhoE_new = ub[3] > e_min ? ub[3] : e_min;
    const double d0 = rho_new - ub[0], d3 = rhoE_new - ub[3];
    if (d0 != 0.0 || d3 != 0.0) {
        for (int i = 0; i < C::N; ++i) {
            u[(i * 4) * stride] += d0;
            u[(i * 4 + 3) * stride] += d3;
        }
        ub[0] = rho_new;
        ub[3] = rhoE_new;
        fl |= 1;
    }
    const double p_bar = g1 * (ub[3] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2]) / ub[0]);
    double eps_rho = P.fraction * ub[0];
    if (eps_rho < 0.5 * P.rho_floor) eps_rho = 0.5 * P.rho_floor;
    double eps_p = P.fraction * p_bar;
    if (eps_p < 0.5 * P.p_floor) eps_p = 0.5 * P.p_floor;
    // O(N) sufficient screen (_numba_kernels.py:1020-1045); never changes the answer
    if (P.lebesgue > 0.0) {
        double dr = 0, dmx = 0, dmy = 0, den = 0;
        for (int i = 0; i < C::N; ++i) {
            dr = fmax(dr, fabs(u[(i * 4) * stride] - ub[0]));
            dmx = fmax(dmx, fabs(u[(i * 4 + 1) * stride] - ub[1]));
            dmy = fmax(dmy, fabs(u[(i * 4 + 2) * stride] - ub[2]));
            den = fmax(den, fabs(u[(i * 4 + 3) * stride] - ub[3]));
        }
        const double rho_lo = ub[0] - P.lebesgue * dr;
        if (rho_lo >= eps_rho) {
            const double m_hi =
                sqrt(ub[1] * ub[1] + ub[2] * ub[2]) + P.lebesgue * sqrt(dmx * dmx + dmy * dmy);
            const double p_lo = g1 * (ub[3] - P.lebesgue * den - 0.5 * m_hi * m_hi / rho_lo);
            if (p_lo >= eps_p) return fl;
        }
    }
    double pmin;
    const double rmin = probe_min<C>(u, stride, ub, 1.0, tb, side, g, &pmin);
    if (rmin >= eps_rho && pmin >= eps_p) return fl;
    fl |= 2;
    double theta = 1.0;
    if (rmin < eps_rho) {
        const double gap = ub[0] - rmin;
        theta = gap <= 0.0 ? 0.0 : fmin(fmax((ub[0] - eps_rho) / gap, 0.0), 1.0);
    }
    double phi;
    probe_min<C>(u, stride, ub, theta, tb, side, g, &phi);
    if (phi < eps_p) {
        double lo = 0.0, hi = theta;
#pragma unroll 1
        for (int it = 0; it < P.bisect_steps; ++it) {
            const double mid = 0.5 * (lo + hi);
            double pm;
            probe_min<C>(u, stride, ub, mid, tb, side, g, &pm);
            if (pm >= eps_p) lo = mid;
            else hi = mid;
        }
        theta = lo;
    }
    for (int i = 0; i < C::N; ++i)
        for (int s = 0; s < 4; ++s) {
            double* x = &u[(i * 4 + s) * stride];
            *x = ub[s] + theta * (*x - ub[s]);
        }
    *scaled = 1;
    return fl;
}

__device__ __forceinline__ int wave_slot() {
    return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int t = threadIdx.x;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double tot = 0.0;
    if (t == 0)
        for (int w = 0; w < BS / 64; ++w) tot += red[w];
    return tot;
}

// ---------------------------------------------------------------- DG kernel
// Component-split mapping: a workgroup of one wave owns 16 consecutive
// elements and gives each element four lanes, one per conserved variable
// (lane = 4 * element + s).  Everything linear in U -- traces, volume-point
// interpolation, lifting, volume projection, the stage combination and the
// limiter's probe sums -- is separable by component, so each lane keeps its
// element's coefficient column U[:, s] and its accumulator R[:, s] in
// registers and reads reference tables with scalar loads.  Only the
// pointwise physics needs all four components; it runs on (point, element)
// items, one per lane, exchanging point states through LDS rows
// [point][component][element].  The limiter gathers the other components of
// its element with quad shuffles.
//
// Face traces are formed in the face's canonical point order (the side-0
// table: sigma increasing along the local face), so all lanes use one table.
// Gauss-Legendre points are symmetric, so at edge point q the side-0 element
// reads its canonical index q and the side-1 element nqf-1-q; the flux item
// writes the weighted flux back at the element's canonical index and the
// lift is side-independent as well.
constexpr int DG_BS = 64;  // one wave per workgroup: no cross-wave barriers
using d2 = double __attribute__((ext_vector_type(2)));  // a 16-byte pair of doubles

// A copy of x that the compiler cannot prove equal to x: indices derived from
// it are recomputed where they are used instead of being kept live (or
// spilled) from an earlier computation of the same expression.
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}
constexpr int TE = 16;     // elements per wave tile (four lanes each)
constexpr int QS = 80;     // LDS doubles per point: four component rows of TE, = 16 mod 64

// LDS slot of component c of tile element el at point q.  Rows are TE wide;
// the XOR on the element index and the point stride keep both access
// patterns free of bank conflicts: four lanes per element (lane = 4 el + c)
// and one lane per (point, element) (lane = TE q + el, c fixed).
__device__ __forceinline__ int at(int q, int c, int el) {
    return q * QS + c * TE + (el ^ ((c >> 1) << 3));
}

// The per-element inverse mass matrices of non-affine elements are stored as
// their upper triangles, row-major (N (N + 1) / 2 doubles each, 360 instead of
// 648 bytes per element at N = 9): entry (i, j) and (j, i) is the upper one.
__host__ __device__ constexpr int sym_size(int n) { return n * (n + 1) / 2; }
__host__ __device__ constexpr int sym_index(int n, int i, int j) {
    return i <= j ? i * n - i * (i - 1) / 2 + (j - i) : j * n - j * (j - 1) / 2 + (i - j);
}

template <class C>
struct Tile {
    static constexpr int N = C::N, NQV = C::NQV, NQF = C::NQF, NF = C::NF, NS = C::NS;
    static constexpr int QR = DG_BS / TE;  // volume points per round: one per lane
    // own trace / weighted flux, neighbour trace, signal speeds.  One buffer
    // each: a one-wave workgroup executes its LDS operations in program
    // order, so the next face may overwrite rows the last lift has read.
    static constexpr int FACE = 2 * NQF * QS + NQF * TE;
    // point states / xi-flux, eta-flux
    static constexpr int VOL = 2 * QR * QS;
    // staging of U in and out: the tile's rows exactly as in HBM, so the
    // copies in and out are plain contiguous moves
    static constexpr int SIO = NS;
    static constexpr int IO = TE * SIO;
    static constexpr int WORK = FACE > VOL ? (FACE > IO ? FACE : IO) : (VOL > IO ? VOL : IO);
    // staged once per tile: edge geometry of every element face (nx, ny,
    // |dX/dsigma| per edge, or per edge point when curved) and the affine
    // metrics
    static constexpr int GF = C::STR ? 3 : 3 * NQF;
    static constexpr int EGN = TE * NF * GF;
    static constexpr int GEN = C::AFF ? TE * 5 : 1;
    // LDS bytes of one tile; when four tiles fit each SIMD's share of the
    // CU's 160 KB, the register budget is capped at four waves per SIMD
    // (128 VGPRs) so that the LDS and the registers allow the same
    // occupancy: the fourth wave hides more of the latency than the few
    // spilled registers cost.
    static constexpr int LDS = 8 * (WORK + NF * N * NQF + EGN + GEN + TE) + 8 * TE * NF;
    // (also when the LDS allows 14 or 15 tiles per CU, where most SIMDs still
    // hold four; P3 quadrilaterals need more registers than four allow)
    static constexpr int WPE = LDS * 14 <= 160 * 1024 && N < 16 ? 4 : 1;
    // a later stage's rows of Un in registers: one double per lane per round
    using Row = double;
    static constexpr int NR = TE * NS / DG_BS;
    __device__ static void put_rows(double* w, int t, const Row (&r)[NR]) {
#pragma unroll
        for (int j = 0; j < NR; ++j) w[t + j * DG_BS] = r[j];
    }
};

// Adds the number of active lanes whose pred holds to *c with one atomic per
// wave.  Same-address atomics from every lane would serialise at the memory
// side (the counter is shared by all XCDs).  Call from converged code.
__device__ __forceinline__ void wave_count(unsigned long long* c, bool pred) {
    const unsigned long long m = __ballot(pred);
    const unsigned long long act = __ballot(true);
    if (m && static_cast<int>(__lane_id()) == __ffsll(static_cast<long long>(act)) - 1)
        atomicAdd(c, static_cast<unsigned long long>(__popcll(m)));
}

// LDS hand-off between the lanes of a one-wave workgroup.  A wave's LDS
// operations complete in program order, so lanes see each other's earlier
// writes without waiting for them; what is needed is that the compiler keep
// the accesses on their side of this point (a wavefront-scope fence), not a
// workgroup barrier, whose wait for every outstanding load and store would
// also stall the global and scalar loads issued ahead of their use.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// component k of this lane's element (lanes 4e..4e+3 hold s = 0..3)
__device__ __forceinline__ double quad_get(double v, int k) {
    return __shfl(v, (static_cast<int>(threadIdx.x) & ~3) | k, 64);
}

// sum_i b(i) col[i] with no contraction: the limiter's probe sums must round
// as the CPU oracle's (src/backends/_numba_kernels.py:878-884)
template <int N, class B>
__device__ __forceinline__ double exact_dot(const double* col, B b) {
#pragma clang fp contract(off)
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) v += b(i) * col[i];
    return v;
}

// The probe set: the volume points and the element's own face points
// (src/backends/_numba_kernels.py:858-924).  probe_value is this lane's
// component at probe k, summed without contraction.
template <class C>
__device__ __forceinline__ double probe_value(const double* col, int k,
                                              const double* __restrict__ T, const nz_problem& P,
                                              int sidebits) {
#pragma clang fp contract(off)
    constexpr int N = C::N, NQV = C::NQV, NQF = C::NQF, NF = C::NF;
    if (k < NQV) return exact_dot<N>(col, [&](int i) { return T[P.off.phi_vol + i * NQV + k]; });
    const int f = (k - NQV) / NQF, q = (k - NQV) - f * NQF;
    const int sd = (sidebits >> f) & 1;
    return exact_dot<N>(col, [&](int i) {
        const double b0 = T[P.off.phi_face + ((0 * NF + f) * N + i) * NQF + q];
        const double b1 = T[P.off.phi_face + ((1 * NF + f) * N + i) * NQF + q];
        return sd ? b1 : b0;
    });
}

template <class C>
struct Probes {
    static constexpr int NP = C::NQV + C::NF * C::NQF;
    // probes per lane: lane s of an element holds the four conserved
    // variables of probes s, s + 4, s + 8, ...
    static constexpr int NK = (NP + 3) / 4;
};

// The probe states of this lane's share, formed once per limited element:
// every lane sums its own component at each probe (no contraction), and the
// lane that owns probe k (k mod 4) collects the four components.  The probe
// loop is not unrolled, so one probe's table column is in scalar registers at
// a time; the store into the share goes through a wave-uniform test per slot.
template <class C>
__device__ __forceinline__ void probe_share(const double* col, int s, const double* __restrict__ T,
                                            const nz_problem& P, int sidebits,
                                            double V[Probes<C>::NK][4]) {
    constexpr int NP = Probes<C>::NP;
#pragma unroll 1
    for (int k = 0; k < NP; ++k) {
        const double v = probe_value<C>(col, k, T, P, sidebits);
        double x[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) x[c] = quad_get(v, c);
        const bool mine = (k & 3) == s;
        const int kk = k >> 2;
#pragma unroll
        for (int j = 0; j < Probes<C>::NK; ++j)
            if (j == kk && mine)
                for (int c = 0; c < 4; ++c) V[j][c] = x[c];
    }
#pragma unroll
    for (int kk = 0; kk < Probes<C>::NK; ++kk)
        if (4 * kk + s >= NP)  // no such probe: repeat one that exists
            for (int c = 0; c < 4; ++c) V[kk][c] = V[0][c];
}

// min density and pressure of ub + theta (V - ub) over the element's probes;
// every lane of the element returns the same values.  Each lane scans its
// share and the four lanes combine with two shuffles; min is exact, so the
// result is the same as a scan over all probes in order.
template <class C>
__device__ __forceinline__ void probe_minmax(const double V[Probes<C>::NK][4], const double ub[4],
                                             double theta, double g, double* rmin, double* pmin) {
#pragma clang fp contract(off)
    double rm = 1e300, pm = 1e300;
#pragma unroll
    for (int kk = 0; kk < Probes<C>::NK; ++kk) {
        double W[4];
        for (int c = 0; c < 4; ++c) W[c] = ub[c] + theta * (V[kk][c] - ub[c]);
        rm = W[0] < rm ? W[0] : rm;
        const double r = W[0] > 1e-300 ? W[0] : 1e-300;
        const double pp = (g - 1.0) * (W[3] - 0.5 * (W[1] * W[1] + W[2] * W[2]) / r);
        pm = pp < pm ? pp : pm;
    }
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
        const double ro = __shfl_xor(rm, o, 64), po = __shfl_xor(pm, o, 64);
        rm = ro < rm ? ro : rm;
        pm = po < pm ? po : pm;
    }
    *rmin = rm;
    *pmin = pm;
}

// Zhang-Shu positivity limiter on the element whose column `col` this lane
// holds (src/backends/_numba_kernels.py:927-1093; flags as
// src/spatial/limiter.py:125-129,222-224).  The four lanes of an element take
// identical branches.  Returns the flag bits; *scaled = 1 when theta < 1 was
// applied.
template <class C>
__device__ uint8_t limit_column(double* col, int s, const double* mw, const double* __restrict__ T,
                                const nz_problem& P, int sidebits, int* scaled) {
#pragma clang fp contract(off)
    constexpr int N = C::N;
    const double g1 = P.gamma - 1.0;
    uint8_t fl = 0;
    double ub[4];
    {
        const double m = exact_dot<N>(col, [&](int i) { return mw[i]; });
        for (int c = 0; c < 4; ++c) ub[c] = quad_get(m, c);
    }
    const double rho_new = ub[0] > P.rho_floor ? ub[0] : P.rho_floor;
    const double e_min = P.p_floor / g1 + 0.5 * (ub[1] * ub[1] + ub[2] * ub[2]) / rho_new;
    const double rhoE_new = ub[3] > e_min ? ub[3] : e_min;
    const double d0 = rho_new - ub[0], d3 = rhoE_new - ub[3];
    if (d0 != 0.0 || d3 != 0.0) {
        const double d = s == 0 ? d0 : (s == 3 ? d3 : 0.0);
        if (s == 0 || s == 3)
            for (int i = 0; i < N; ++i) col[i] += d;
        ub[0] = rho_new;
        ub[3] = rhoE_new;
        fl |= 1;
    }
    const double p_bar = g1 * (ub[3] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2]) / ub[0]);
    double eps_rho = P.fraction * ub[0];
    if (eps_rho < 0.5 * P.rho_floor) eps_rho = 0.5 * P.rho_floor;
    double eps_p = P.fraction * p_bar;
    if (eps_p < 0.5 * P.p_floor) eps_p = 0.5 * P.p_floor;
    if (P.lebesgue > 0.0) {  // O(N) sufficient screen (_numba_kernels.py:1020-1045)
        double dm = 0.0;
        for (int i = 0; i < N; ++i) dm = fmax(dm, fabs(col[i] - ub[s]));
        double dv[4];
        for (int c = 0; c < 4; ++c) dv[c] = quad_get(dm, c);
        const double rho_lo = ub[0] - P.lebesgue * dv[0];
        if (rho_lo >= eps_rho) {
            const double m_hi = sqrt(ub[1] * ub[1] + ub[2] * ub[2]) +
                                P.lebesgue * sqrt(dv[1] * dv[1] + dv[2] * dv[2]);
            const double p_lo = g1 * (ub[3] - P.lebesgue * dv[3] - 0.5 * m_hi * m_hi / rho_lo);
            if (p_lo >= eps_p) return fl;
        }
    }
    double V[Probes<C>::NK][4];
    probe_share<C>(col, s, T, P, sidebits, V);
    double rmin, pmin;
    probe_minmax<C>(V, ub, 1.0, P.gamma, &rmin, &pmin);
    if (rmin >= eps_rho && pmin >= eps_p) return fl;
    fl |= 2;
    double theta = 1.0;
    if (rmin < eps_rho) {
        const double gap = ub[0] - rmin;
        theta = gap <= 0.0 ? 0.0 : fmin(fmax((ub[0] - eps_rho) / gap, 0.0), 1.0);
    }
    double rr, phi;
    probe_minmax<C>(V, ub, theta, P.gamma, &rr, &phi);
    if (phi < eps_p) {
        double lo = 0.0, hi = theta;
#pragma unroll 1
        for (int it = 0; it < P.bisect_steps; ++it) {
            const double mid = 0.5 * (lo + hi);
            double pm;
            probe_minmax<C>(V, ub, mid, P.gamma, &rr, &pm);
            if (pm >= eps_p) lo = mid;
            else hi = mid;
        }
        theta = lo;
    }
    for (int i = 0; i < N; ++i) col[i] = ub[s] + theta * (col[i] - ub[s]);
    *scaled = 1;
    return fl;
}

// the staged tile (rows of NS doubles, as in HBM) back to global memory
template <class L>
__device__ __forceinline__ void store_tile(double* __restrict__ out, const double* tile, int t,
                                           int ne) {
    // streaming stores: the tile is not read again by this launch, so it is
    // kept out of the way of the neighbour gathers in L2
    constexpr int NU = TE * L::NS / DG_BS;
    if (ne == TE) {
#pragma unroll
        for (int j = 0; j < NU; ++j) __builtin_nontemporal_store(tile[t + j * DG_BS], &out[t + j * DG_BS]);
    } else {
        for (int it = t; it < ne * L::NS; it += DG_BS) __builtin_nontemporal_store(tile[it], &out[it]);
    }
}

// The volume terms of an element tile (shared by k_dg and k_dgp): the state
// interpolated to the volume points (four points per round, one per lane
// group), the Euler fluxes at one (point, element) item per lane, their
// contravariant components through LDS, and the projection onto the basis
// accumulated into acc.  eid(pe) is the element held by tile slot pe.
template <class C, class L, class EID>
__device__ __forceinline__ void dg_volume(const nz_problem& P, const double* __restrict__ T,
                                          double* sW, const double* sGe,
                                          const double (&u)[C::N], double (&acc)[C::N], int t,
                                          int cs, int ne, bool fold, EID eid) {
    constexpr int N = C::N, NQV = C::NQV;
    double* sUq = sW;                   // point states, then xi-flux
    double* sFy = sW + L::QR * QS;      // eta-flux
    // point-major volume tables: phi, dxi, deta, axi, aeta (N each) per point
    const double* vt = T + P.off.vol_t;
    const int vproj = fold ? 3 * N : N;
#pragma unroll 1
    for (int q0 = 0; q0 < NQV; q0 += L::QR) {
        // per-point metrics of this lane's (point, element) item: requested
        // before the interpolation, whose FMAs cover their latency
        // (unconditional, at a valid point, so no branch join waits on them)
        double gm[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if constexpr (!C::AFF) {
            const int qq = t / TE, pe = t % TE, q = q0 + qq;
            const bool ok = q < NQV && pe < ne;
            const double* gq = P.geo + (eid(ok ? pe : 0) * NQV + (ok ? q : 0)) * 5;
#pragma unroll
            for (int k = 0; k < 5; ++k) gm[k] = gq[k];
        }
#pragma unroll 2  // two points' table rows in flight
        for (int qq = 0; qq < L::QR; ++qq) {
            if (q0 + qq < NQV) {
                double a = 0.0;
                const double* ph = vt + (q0 + qq) * 5 * N;
#pragma unroll
                for (int i = 0; i < N; ++i) a += ph[i] * u[i];
                sUq[qq * QS + cs] = a;
            }
        }
        wave_sync();
        {
            const int qq = t / TE, pe = t % TE, q = q0 + qq;
            if (q < NQV && pe < ne) {
                double uq[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) uq[c] = sUq[at(qq, c, pe)];
                const Prim<double> w = primitives(uq, P.gamma);
                const double F[4] = {w.r * w.u, w.r * w.u * w.u + w.p, w.r * w.u * w.v, w.r * w.u * w.H};
                const double G[4] = {w.r * w.v, w.r * w.u * w.v, w.r * w.v * w.v + w.p, w.r * w.v * w.H};
                // contravariant metrics; unfolded, they carry the quadrature
                // weight times det J (the folded tables carry it already)
                double gx, gy, hx, hy;
                if (C::AFF) {
                    const double* ge = sGe + pe * 5;
                    gx = ge[0];
                    gy = ge[1];
                    hx = ge[2];
                    hy = ge[3];
                    if (!fold) {
                        const double wd = T[P.off.w_vol + q] * ge[4];
                        gx *= wd;
                        gy *= wd;
                        hx *= wd;
                        hy *= wd;
                    }
                } else {
                    const double wd = gm[4];
                    gx = gm[0] * wd;
                    gy = gm[1] * wd;
                    hx = gm[2] * wd;
                    hy = gm[3] * wd;
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    sUq[at(qq, c, pe)] = gx * F[c] + gy * G[c];
                    sFy[at(qq, c, pe)] = hx * F[c] + hy * G[c];
                }
            }
        }
        wave_sync();
#pragma unroll 1
        for (int qq = 0; qq < L::QR; ++qq) {
            const int q = q0 + qq;
            if (q < NQV) {
                const double a = sUq[qq * QS + cs], b = sFy[qq * QS + cs];
                const double* pr = vt + q * 5 * N + vproj;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    acc[i] = fma(-pr[i], a, acc[i]);
                    acc[i] = fma(-pr[N + i], b, acc[i]);
                }
            }
        }
        // the next round rewrites these rows after the reads above
    }

}

// The boundary fluxes of a stage's output for the next stage (nz_problem
// bflux_chain bit 1), from the output rows in LDS: one (boundary face, point)
// item per lane over the tile's own boundary faces, formed as k_bflux forms
// them, so the next stage reads the same bits the pre-pass would write.  A
// boundary face belongs to one element, so the workgroup that reads its entry
// in this stage is the one that rewrites it, after the read.
template <class C, class L>
__device__ __forceinline__ void boundary_next(const nz_problem& P, const double* __restrict__ T,
                                              const double* sW, const int* sIn, int t, int ne) {
    constexpr int N = C::N, NQF = C::NQF, NF = C::NF, SL = TE * NF;
    static_assert(SL <= DG_BS, "one lane per face slot");
    __shared__ uint8_t sBl[SL];  // the tile's boundary face slots, in slot order
    const bool bnd = t < SL && t / NF < ne && (sIn[t < SL ? t : 0] >> 3) >= P.n_int;
    const unsigned long long m = __ballot(bnd);
    if (m == 0ull) return;
    if (bnd) sBl[__popcll(m & ((1ull << t) - 1ull))] = static_cast<uint8_t>(t);
    wave_sync();
    const int n_it = __popcll(m) * NQF;
    for (int it = t; it < n_it; it += DG_BS) {
        const int j = it / NQF, q = it - j * NQF;
        const int sl = sBl[j], el = sl / NF, f = sl - el * NF;
        const int64_t k = sIn[sl] >> 3, b = k - P.n_int;
        const int tag = P.bedge[2 * b + 1] & 7;
        double uo[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < N; ++i) {
            const double w = T[P.off.phi_face + (f * N + i) * NQF + q];
#pragma unroll
            for (int c = 0; c < 4; ++c) uo[c] += w * sW[el * L::SIO + i * 4 + c];
        }
        const double* eg = C::STR ? P.edge_geo + k * 3 : P.edge_geo + (k * NQF + q) * 3;
        double fx[4];
        const double sp = boundary_flux(P, tag, uo, eg[0], eg[1], P.p_back, fx);
        const double sc = eg[2] * T[P.off.w_face + q];
        double* out = P.bflux + (b * NQF + q) * 5;
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = fx[c] * sc;
        out[4] = sp;
    }
}

// The epilogue of an element tile (shared by k_dg and k_dgp): the inverse
// mass of non-affine elements, then either the residual / rate rows, or the
// fused Runge-Kutta stage (time step, combination, norm partial, positivity
// limiter); the tile's rows staged in LDS go out through store().
template <class C, class L, bool STAGE, class STORE, class LOADN>
__device__ __forceinline__ void dg_epilogue(
    const nz_problem& P, const double* __restrict__ T, double* sW, const int* sIn,
    double (&u)[C::N], double (&acc)[C::N], double wave, int t, int el, int s, int64_t e, bool live,
    int ne, int64_t slot, double* __restrict__ dt, double dt_factor, double ca, double cb, int mode,
    double* __restrict__ partial, uint8_t* __restrict__ flags,
    unsigned long long* __restrict__ counters, STORE store, LOADN load_un) {
    constexpr int N = C::N, NF = C::NF;
    // a later stage combines with Un: its rows are requested first, in the
    // tile's row layout (coalesced), and handed to the lanes through LDS
    // (non-affine elements request them after the inverse mass, whose rows
    // need the registers)
    typename L::Row rn[L::NR];
    if (STAGE && mode == kStageNext && C::AFF) load_un(rn);
    // acc holds R (unfolded) or -rate (affine, folded) for component s
    if ((STAGE || mode != kRes) && !C::AFF) {
        // (idle lanes read element 0's matrix and discard the products)
        const double* me = P.minv + (live ? e : 0) * sym_size(N);
        if constexpr (N >= 9) {
            // Large matrices: two rows of M^-1 in flight at a time (row i + 1
            // is requested before row i is summed; the fence keeps row i + 2
            // behind it), the products parked in this lane's LDS row.  The
            // sums are the unrolled form's, term for term.
            wave_sync();  // the volume rows' last reads come first
            double m[2][N];
#pragma unroll
            for (int j = 0; j < N; ++j) m[0][j] = me[sym_index(N, 0, j)];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (i + 1 < N) {
#pragma unroll
                    for (int j = 0; j < N; ++j) m[(i + 1) & 1][j] = me[sym_index(N, i + 1, j)];
                }
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < N; ++j) a += m[i & 1][j] * acc[j];
                sW[el * L::SIO + i * 4 + s] = a;
                asm volatile("" ::: "memory");
            }
#pragma unroll
            for (int i = 0; i < N; ++i) acc[i] = sW[el * L::SIO + i * 4 + s];
        } else if (live) {
            double r[N];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < N; ++j) a += me[sym_index(N, i, j)] * acc[j];
                r[i] = a;
            }
#pragma unroll
            for (int i = 0; i < N; ++i) acc[i] = r[i];
        }
    }
    if (STAGE && mode == kStageNext && !C::AFF) load_un(rn);
    wave_sync();  // last reads of the volume rows are done; reuse sW for staging
    if (!STAGE) {
        const double osign = mode == kRes ? 1.0 : -1.0;  // rate = -M^-1 R
#pragma unroll
        for (int i = 0; i < N; ++i) sW[el * L::SIO + i * 4 + s] = osign * acc[i];
        if (s == 0 && live) dt[e] = wave;
        wave_sync();
        store();
        return;
    }

    // fused stage: time step, combination, norm partials, limiter, store
    if (mode == kStageNext) {
        L::put_rows(sW, t, rn);
        wave_sync();
    }
    double h = 0.0;
    if (live) {
        if (mode == kStageFirst) {
            const double d = wave < kFloor ? kFloor : wave;
            h = dt_factor * 2.0 * P.area[e] / d;
            if (s == 0) dt[e] = h;
        } else {
            h = dt[e];
        }
    }
    double nacc = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double F = -acc[i];
        if (mode == kStageFirst) {
            u[i] = u[i] + h * F;
            const double v = F * P.wnorm[s];
            nacc += live ? v * v : 0.0;
        } else {
            u[i] = live ? ca * sW[el * L::SIO + i * 4 + s] + cb * (u[i] + h * F) : 0.0;
        }
    }
    if (mode == kStageFirst) {
        double tot = nacc;
        for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (t == 0) partial[slot] = tot;
    }
    if (P.limiter == 1 && N > 1) {  // superbee runs as separate passes
        uint8_t fl = 0;
        int scaled = 0;
        if (live) {
            int sidebits = 0;
#pragma unroll
            for (int f = 0; f < NF; ++f) sidebits |= ((sIn[el * NF + f] >> 2) & 1) << f;
            double mw[N];
#pragma unroll
            for (int i = 0; i < N; ++i) mw[i] = C::AFF ? T[P.off.mean_w + i] : P.mean_w[e * N + i];
            fl = limit_column<C>(u, s, mw, T, P, sidebits, &scaled);
        }
        if (counters) {
            wave_count(&counters[0], live && s == 0 && scaled);
            wave_count(&counters[1], live && s == 0 && (fl & 1));
        }
        if (flags && live && s == 0) flags[e] = fl;
    } else if (flags && live && s == 0) {
        flags[e] = 0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) sW[el * L::SIO + i * 4 + s] = u[i];
    wave_sync();
    store();
    // (stages only: the residual and rate entry points refuse the flag)
    if (STAGE && (P.bflux_chain & 2)) boundary_next<C, L>(P, T, sW, sIn, t, ne);
}

template <class C, int FLUX, bool STAGE>
__global__ __launch_bounds__(DG_BS) __attribute__((amdgpu_waves_per_eu(Tile<C>::WPE)))
void k_dg(const nz_problem P, const double* __restrict__ T,
                                           const double* __restrict__ Uin,
                                           const double* __restrict__ Un,
                                           double* __restrict__ Uout, double* __restrict__ dt,
                                           double dt_factor, double ca, double cb, int mode,
                                           double* __restrict__ partial,
                                           uint8_t* __restrict__ flags,
                                           unsigned long long* __restrict__ counters) {
    using L = Tile<C>;
    constexpr int N = C::N, NQV = C::NQV, NQF = C::NQF, NF = C::NF, NS = C::NS;
    constexpr int NPF = NF * N * NQF;  // canonical (side-0) face table
    constexpr int NI = C::SPT ? C::FN : N;  // neighbour coefficients a face trace reads
    __shared__ double sW[L::WORK];
    __shared__ double sPF[NPF];
    __shared__ double sEG[L::EGN];
    __shared__ double sGe[L::GEN];
    __shared__ double sRd[C::AFF ? TE : 1];  // 1 / det J of the affine elements
    __shared__ int sNb[TE * NF], sIn[TE * NF];

    const int t = threadIdx.x;
    const int el = t >> 2, s = t & 3;  // component lanes
    int64_t e0, slot;
    int ne;
    if (P.tiles) {  // tile table: dispatch order, disjoint meshes, or a subset of tiles
        e0 = P.tiles[3 * blockIdx.x];
        ne = static_cast<int>(P.tiles[3 * blockIdx.x + 1] - e0);
        slot = P.tiles[3 * blockIdx.x + 2];
    } else {
        e0 = static_cast<int64_t>(blockIdx.x) * TE;
        ne = static_cast<int>(P.n_elem - e0 < TE ? P.n_elem - e0 : TE);
        slot = blockIdx.x;
    }
    const bool live = el < ne;
    const int64_t e = e0 + el;
    const bool fold = C::AFF && (STAGE || mode != kRes);  // affine rate: M^-1 folded into tables
    const int cs = at(0, s, el);  // this lane's slot at point 0

    // A wave staging its tile runs at raised priority, so its loads are issued
    // ahead of the other waves' arithmetic and their latency overlaps it.
    __builtin_amdgcn_s_setprio(2);
    // ---- stage the tile (coalesced), the face maps, the face table and the
    // affine metrics; every load is issued before the first LDS write
    {
        // Every load is unconditional (entries past the tile read a valid
        // address and are masked: loads under branches would end in a full
        // wait where the branches join), and the edge geometry is requested
        // with the rest: lane r holds element-face r's face map (a tile has
        // at most 64 faces).
        static_assert(TE * NF <= DG_BS, "one lane per element face");
        constexpr int NU = TE * NS / DG_BS;  // = N
        constexpr int NT = (NPF + DG_BS - 1) / DG_BS;
        constexpr int NG = (L::GEN + DG_BS - 1) / DG_BS;
        constexpr int NE_ = (L::EGN + DG_BS - 1) / DG_BS;
        double ru[NU], rt[NT], rg[NG], rq[NE_];
        int rn, ri;
        {
            const bool ok = t < ne * NF;
            const int a = P.fe_nbr[e0 * NF + (ok ? t : 0)], b = P.fe_info[e0 * NF + (ok ? t : 0)];
            rn = ok ? a : -1;
            ri = ok ? b : 0;
        }
#pragma unroll
        for (int j = 0; j < NE_; ++j) {
            const int it = t + j * DG_BS;
            const int r = it / L::GF, jj = it - r * L::GF;  // r = element-face
            const bool ok = it < L::EGN && r < ne * NF;
            const int64_t k = __shfl(ri, ok ? r : 0, 64) >> 3;
            const double v = P.edge_geo[(ok ? k : 0) * L::GF + (ok ? jj : 0)];
            rq[j] = ok ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int it = t + j * DG_BS, r = it / NS;
            const double v = Uin[e0 * NS + (r < ne ? it : 0)];
            ru[j] = r < ne ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int it = t + j * DG_BS;
            const double v = T[P.off.phi_face + (it < NPF ? it : 0)];
            rt[j] = it < NPF ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            const int it = t + j * DG_BS;
            const bool ok = C::AFF && it < ne * 5;
            const double v = C::AFF ? P.geo[e0 * 5 + (ok ? it : 0)] : 0.0;
            rg[j] = ok ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) sW[t + j * DG_BS] = ru[j];
        if (t < TE * NF) {
            sNb[t] = rn;
            sIn[t] = ri;
        }
#pragma unroll
        for (int j = 0; j < NT; ++j)
            if (t + j * DG_BS < NPF) sPF[t + j * DG_BS] = rt[j];
#pragma unroll
        for (int j = 0; j < NG; ++j)
            if (t + j * DG_BS < L::GEN) sGe[t + j * DG_BS] = rg[j];
#pragma unroll
        for (int j = 0; j < NE_; ++j)
            if (t + j * DG_BS < L::EGN) sEG[t + j * DG_BS] = rq[j];
    }
    double u[N], acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u[i] = sW[el * L::SIO + i * 4 + s];
        acc[i] = 0.0;
    }
    __builtin_amdgcn_s_setprio(0);
    // folded lift signs: (+-1) / det J, as +-(1 / det J) (the same double)
    if (C::AFF && t < TE) sRd[t] = 1.0 / sGe[t * 5 + 4];
    double wave = 0.0;
    wave_sync();  // the staging rows are reused below

    double* sTo = sW;                 // own trace, then weighted flux
    double* sTn = sW + NQF * QS;      // neighbour trace
    double* sSp = sW + 2 * NQF * QS;  // [NQF][TE] signal speed
    const int lift_off = fold ? P.off.lift : P.off.phi_face;

    // ---------------------------------------------------------------- faces
#pragma unroll 1
    for (int f = 0; f < NF; ++f) {
        const int nb = sNb[el * NF + f], inf = sIn[el * NF + f];
        const bool has = live && nb >= 0;
        const int key = has ? (inf & 3) : -1;  // the neighbour's local face
        // the neighbour's column of component s at the nodes of its face,
        // requested before the own trace is formed (lanes without a
        // neighbour read the tile's first element and discard it)
        double xn[N];
        {
            const double* ur = Uin + (has ? static_cast<int64_t>(nb) : e0) * NS + s;
            const int kk = has ? key : 0;
#pragma unroll
            for (int j = 0; j < NI; ++j) xn[j] = ur[nbr_node<C>(kk, j) * 4];
        }
        // (e,s): own and neighbour traces of component s, canonical order
        {
            const double (*cols[1])[N] = {&u};
            trace_face_rt<C, 1>(f, T + P.off.phi_face, cols,
                                [&](int q, const double* a) { sTo[q * QS + cs] = a[0]; });
        }
        {
            // The neighbour's face table is read with scalar loads when the
            // wave agrees on it (the usual case on a structured mesh) and
            // from LDS otherwise.
            // (rolled point loops write each point's trace straight to LDS)
            double tn[NQF];
#pragma unroll
            for (int q = 0; q < NQF; ++q) tn[q] = 0.0;
            const int key0 = __builtin_amdgcn_readfirstlane(key);
            const double (*cn[1])[N] = {&xn};
            auto put = [&](int q, const double* a) {
                if constexpr (C::ROLL_T) sTn[q * QS + cs] = a[0];
                else tn[q] = a[0];
            };
            if (__all(key == key0) && key0 >= 0) {
                const double* ph = T + P.off.phi_face + key0 * N * NQF;
                trace_nbr<C, 1>(key0, [&](int k) { return ph[k]; }, cn, put);
            } else if (has) {
                const double* ph = sPF + key * N * NQF;
                trace_nbr<C, 1>(key, [&](int k) { return ph[k]; }, cn, put);
            } else if constexpr (C::ROLL_T) {
#pragma unroll
                for (int q = 0; q < NQF; ++q) sTn[q * QS + cs] = 0.0;
            }
            if constexpr (!C::ROLL_T) {
#pragma unroll
                for (int q = 0; q < NQF; ++q) sTn[q * QS + cs] = tn[q];
            }
        }
        wave_sync();
        // (q,e): the flux at edge point q, one item per lane
        for (int it = t; it < NQF * TE; it += DG_BS) {
            const int q = it / TE, pe = it % TE;
            if (pe < ne) {
                const int nbp = sNb[pe * NF + f], infp = sIn[pe * NF + f];
                const int sd = (infp >> 2) & 1;
                const int64_t k = infp >> 3;
                const int qo = sd ? NQF - 1 - q : q;  // this element's canonical index
                double fx[4], sp;
                if (nbp >= 0) {
                    const int qn = NQF - 1 - qo;      // the neighbour's
                    double ul[4], ur[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const double a = sTo[at(qo, c, pe)];
                        const double b = sTn[at(qn, c, pe)];
                        ul[c] = sd == 0 ? a : b;
                        ur[c] = sd == 0 ? b : a;
                    }
                    const double* eg = sEG + (pe * NF + f) * L::GF + (C::STR ? 0 : q * 3);
                    const double nx = eg[0], ny = eg[1];
                    sp = interface_flux<FLUX>(P, ul, ur, nx, ny, fx);
                    const double sc = eg[2] * face_weight<C>(T, P, q);
#pragma unroll
                    for (int c = 0; c < 4; ++c) fx[c] = fx[c] * sc;
                } else {  // boundary faces: weighted flux precomputed by k_bflux
                    const double* bf = P.bflux + ((k - P.n_int) * NQF + q) * 5;
#pragma unroll
                    for (int c = 0; c < 4; ++c) fx[c] = bf[c];
                    sp = bf[4];
                }
                double sg = sd ? -1.0 : 1.0;
                if (fold) sg = sd ? -sRd[pe] : sRd[pe];
#pragma unroll
                for (int c = 0; c < 4; ++c) sTo[at(qo, c, pe)] = sg * fx[c];
                sSp[q * TE + pe] = sp;
            }
        }
        wave_sync();
        // (e,s): lift, and the element's wave-speed sum
        {
            const double* lb = T + lift_off + f * N * NQF;
            if constexpr (C::ROLL_L) {
                lift_face_rolled<C>(lb, [&](int q) { return sTo[q * QS + cs]; }, acc);
            } else {
                double fw[NQF];
#pragma unroll
                for (int q = 0; q < NQF; ++q) fw[q] = sTo[q * QS + cs];
#pragma unroll
                for (int i = 0; i < N; ++i) {
#pragma unroll
                    for (int q = 0; q < NQF; ++q) acc[i] += lb[i * NQF + q] * fw[q];
                }
            }
            if (live) {
                const double* eg = sEG + (el * NF + f) * L::GF;
                double smax = 0.0, len = 0.0;
                for (int q = 0; q < NQF; ++q) {
                    const double sp = sSp[q * TE + el];
                    smax = sp > smax ? sp : smax;
                    len += eg[C::STR ? 2 : q * 3 + 2] * T[P.off.w_face + q];
                }
                wave += smax * len;
            }
        }
        // the next face rewrites these rows after the reads above (one wave:
        // LDS operations complete in program order)
    }
    wave_sync();

    dg_volume<C, L>(P, T, sW, sGe, u, acc, t, cs, ne, fold,
                    [&](int pe) -> int64_t { return e0 + pe; });
    dg_epilogue<C, L, STAGE>(P, T, sW, sIn, u, acc, wave, t, el, s, e, live, ne, slot, dt,
                             dt_factor, ca, cb, mode, partial, flags, counters,
                             [&]() { store_tile<L>(Uout + e0 * NS, sW, t, ne); },
                             [&](typename L::Row (&rn)[L::NR]) {  // the tile's rows of Un
                                 constexpr int NU = L::NR;
                                 const double* src = Un + e0 * NS;
#pragma unroll
                                 for (int j = 0; j < NU; ++j) {
                                     const int x = t + j * DG_BS;
                                     rn[j] = x < ne * NS ? src[x] : 0.0;
                                 }
                             });
}

// ------------------------------------------------------ patch tiles
// k_dgp: k_dg on patch tiles, for affine triangles.  A patch is 16 elements
// that are close in space (a block of 4 x 2 cells of a structured grid, or a
// run of the Morton curve of the element centroids), listed in the patch
// table by element id; 36 of the 48 element faces of a 4 x 2 cell block are
// shared by two of its own elements.  Each interior
// edge's Riemann flux is then solved once, by one (edge, point) item, and
// written into both elements' faces; only faces whose neighbour lies outside
// the patch gather a neighbour column.
//
//   plan    one lane per face slot (element el, face f): the neighbour is in
//           the patch (its slot is the partner), outside it (an off-tile
//           face, numbered in the compact list of those), or the boundary.
//           The edge is owned by its side-0 slot, or by the only slot in the
//           patch; owners are numbered in the item list (wave ballots).
//   traces  every element's traces on all its faces go to LDS, in the
//           face's canonical point order (side-0 table), as k_dg forms them.
//   off     the off-tile neighbours' traces: one (off-tile face, component)
//           item per lane, the neighbour column gathered from global memory.
//   flux    one (edge, point) item per lane: the owner's trace and the
//           partner's, the off-tile trace or the precomputed boundary flux;
//           the weighted flux is written back, signed and folded, into the
//           owner's face and the partner's.  Signal speeds are combined per
//           face with an LDS integer max of their bit patterns (exact for
//           non-negative doubles; NaN speeds are dropped, as in k_dg).
//   lift    each element lifts its faces in face order, as k_dg does.
// Every flux sees exactly the operands k_dg gives it and every sum keeps
// k_dg's order, so the rate and the stage are bit-identical to k_dg's.
constexpr int PATCH_INTS = 18;  // per patch: 16 element ids (-1 padded), count, norm slot
constexpr int XMAX = 12;        // off-tile faces with LDS ghost rows (a 4 x 2 cell block has 12)
constexpr int TRS = 66;         // doubles per face-point row of traces: 16 elements x 4, + 2

// LDS offset of the four components of patch element el in face-point row r
// (rows (face, point)).  Components are contiguous, so a flux item reads an
// element's state with one address; the row stride (66 = 2 mod 32 doubles)
// moves the four points of one item onto distinct banks, and one row is
// written and read by 64 lanes on 64 consecutive doubles.
__device__ __forceinline__ int atq(int r, int el) { return r * TRS + el * 4; }

template <class C>
struct PTile {
    static constexpr int N = C::N, NQF = C::NQF, NF = C::NF, NS = C::NS;
    static constexpr int QR = DG_BS / TE;
    static constexpr int SL = TE * NF;           // face slots of a patch
    static constexpr int TR = NF * NQF * TRS;    // traces, then weighted fluxes: rows (f, q)
    static constexpr int GH = XMAX * NQF * 4;    // off-tile neighbour traces [k][q][c]
    static constexpr int FACE = TR + GH;
    static constexpr int VOL = 2 * QR * QS;
    static constexpr int SIO = NS;
    static constexpr int IO = TE * SIO;
    static constexpr int WORK = FACE > VOL ? (FACE > IO ? FACE : IO) : (VOL > IO ? VOL : IO);
    static constexpr int GEN = TE * 5;
    static constexpr int LDS = 8 * (WORK + GEN + TE + 2 * SL + NQF) + 4 * (2 * SL + TE + XMAX) + 2 * SL;
    // waves per SIMD that the LDS allows, at most four (128 VGPRs)
    static constexpr int WPE = LDS * 16 <= 160 * 1024 ? 4 : (LDS * 12 <= 160 * 1024 ? 3 : 2);
    // The rows move as 16-byte pairs of doubles: an element's NS doubles are
    // an even count, so no pair straddles two elements, and half as many
    // loads and stores need an element id and an address.
    using Row = d2;
    static constexpr int EP = NS / 2;                           // pairs per element
    static constexpr int NR = (TE * EP + DG_BS - 1) / DG_BS;    // pairs per lane
    __device__ static void put_rows(double* w, int t, const Row (&r)[NR]) {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if ((TE * EP) % DG_BS == 0 || t + j * DG_BS < TE * EP)
                reinterpret_cast<Row*>(w)[t + j * DG_BS] = r[j];
    }
};

template <class C, int FLUX, bool STAGE>
__global__ __launch_bounds__(DG_BS) __attribute__((amdgpu_waves_per_eu(PTile<C>::WPE)))
void k_dgp(const nz_problem P, const double* __restrict__ T, const double* __restrict__ Uin,
           const double* __restrict__ Un, double* __restrict__ Uout, double* __restrict__ dt,
           double dt_factor, double ca, double cb, int mode, double* __restrict__ partial,
           uint8_t* __restrict__ flags, unsigned long long* __restrict__ counters) {
    using L = PTile<C>;
    constexpr int N = C::N, NQF = C::NQF, NF = C::NF, NS = C::NS, SL = L::SL;
    constexpr int NI = C::SPT ? C::FN : N;  // neighbour coefficients a face trace reads
    static_assert(C::AFF && C::STR, "patch tiles serve affine elements");
    static_assert(L::SL <= DG_BS, "one lane per face slot");
    static_assert(4 * XMAX <= DG_BS, "one ghost item per lane");
    __shared__ __align__(16) double sW[L::WORK];
    __shared__ double sGe[L::GEN];
    __shared__ double sRd[TE];               // 1 / det J
    __shared__ unsigned long long sSp[SL];   // per face slot: max signal speed (bits)
    __shared__ double sLen[SL];              // per face slot: |dX/dsigma| sum_q w_q
    __shared__ int sNb[SL], sIn[SL], sExt[XMAX], sId[TE];
    __shared__ int8_t sPlan[SL];             // partner slot, -1 boundary, -2 - k off-tile face k
    __shared__ uint8_t sItem[SL];            // flux-owning slots, in slot order
    __shared__ double sWq[NQF];              // face quadrature weights

    const int t = threadIdx.x;
    const int el = t >> 2, s = t & 3;
    const int32_t* pt = P.patches + static_cast<int64_t>(blockIdx.x) * PATCH_INTS;
    const int ne = pt[TE];
    const int64_t slot = pt[TE + 1];
    const bool live = el < ne;
    const int myid = pt[el];
    const int64_t e = live ? myid : 0;  // idle lanes read element 0 and discard it
    const bool fold = STAGE || mode != kRes;
    const int cs = at(0, s, el);  // the volume rows keep k_dg's layout

    // ---- the patch's rows of U, the face maps (one lane per face slot) and
    // the affine metrics, all requested before any is used.  The rows are
    // read as one virtual array of TE rows of NS doubles: the entry at tile
    // offset x belongs to patch element x / NS, so each load instruction
    // covers at most three contiguous runs of memory (an element's NS
    // doubles are contiguous), not one 32-byte piece per element.
    // Every address below comes from the ids in registers (lane 4 r holds
    // element r's), so the face maps and metrics are requested first, then
    // the rows; the edge geometry waits only for the face maps (a wave's
    // vector loads complete in the order they were issued).  The plan needs
    // only the face maps, so it runs, and the off-tile neighbour columns are
    // requested, while the rows are still on their way.
    __builtin_amdgcn_s_setprio(2);
    double u[N], acc[N];
    double g_nx, g_ny, g_jac;
    typename L::Row ru[L::NR];
    int rn, ri;
    double wq;  // face quadrature weight t (t < NQF), for sWq
    {
        constexpr int NG = (L::GEN + DG_BS - 1) / DG_BS;
        // The face maps and metrics are unconditional (a missing element
        // reads the patch's first element, then is masked): loads under
        // branches end in a full wait where the branches join.
        {
            const int r = t / NF;
            const bool ok = t < SL && r < ne;
            const int id = __shfl(myid, 4 * (ok ? r : 0), 64);
            const int64_t o = static_cast<int64_t>(id) * NF + (ok ? t - r * NF : 0);
            const int a = P.fe_nbr[o], b = P.fe_info[o];
            rn = ok ? a : -1;
            ri = ok ? b : 0;
        }
        double rg[NG];
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            const int it = t + j * DG_BS, r = it / 5;
            const bool ok = it < L::GEN && r < ne;
            const int id = __shfl(myid, 4 * (ok ? r : 0), 64);
            const double v = P.geo[static_cast<int64_t>(id) * 5 + (ok ? it - 5 * r : 0)];
            rg[j] = ok ? v : 0.0;
        }
        // this slot's edge geometry (nx, ny, |dX/dsigma|), kept in the slot
        // lane's registers for the flux items to read across lanes
        {
            const double* eg = P.edge_geo + static_cast<int64_t>(ri >> 3) * 3;
            g_nx = eg[0];
            g_ny = eg[1];
            g_jac = eg[2];
        }
        wq = T[P.off.w_face + (t < NQF ? t : 0)];
        // the rows last, unconditionally too: the entries of missing
        // elements repeat the patch's first element, a valid state that only
        // idle lanes compute on and that is never stored
#pragma unroll
        for (int j = 0; j < L::NR; ++j) {
            const int x = t + j * DG_BS, r = x / L::EP;
            const int id = __shfl(myid, 4 * (r < ne ? r : 0), 64);
            const d2* src = reinterpret_cast<const d2*>(Uin + static_cast<int64_t>(id) * NS);
            ru[j] = src[x - r * L::EP];
        }
        if (s == 0) sId[el] = live ? myid : -1;
        if (t < SL) {
            sNb[t] = rn;
            sIn[t] = ri;
        }
#pragma unroll
        for (int j = 0; j < NG; ++j)
            if (t + j * DG_BS < L::GEN) sGe[t + j * DG_BS] = rg[j];
    }
    wave_sync();

    // ---- plan: one lane per face slot
    bool owner = false, offt = false;
    int plan = -1;
    if (t < SL && t / NF < ne) {
        if (rn < 0) {
            owner = true;  // boundary edge: this slot is its side 0
        } else {
            int j = -1;
#pragma unroll
            for (int k = 0; k < TE; ++k) j = sId[k] == rn ? k : j;
            if (j >= 0) {
                plan = j * NF + (ri & 3);
                owner = ((ri >> 2) & 1) == 0;
            } else {
                offt = owner = true;
            }
        }
    }
    const unsigned long long m_own = __ballot(owner), m_off = __ballot(offt);
    const int n_items = __popcll(m_own), n_off = __popcll(m_off);
    {
        const unsigned long long below = (1ull << t) - 1ull;
        if (owner) sItem[__popcll(m_own & below)] = static_cast<uint8_t>(t);
        if (offt) {
            const int k = __popcll(m_off & below);
            if (k < XMAX) sExt[k] = t;
            plan = -2 - k;
        }
        if (t < SL) {
            sPlan[t] = static_cast<int8_t>(plan);
            sSp[t] = 0ull;
        }
        if (t < NQF) sWq[t] = wq;
    }
    wave_sync();

    double* sT = sW;             // traces, then weighted fluxes
    double* sG = sW + L::TR;     // off-tile neighbour traces [k][q][c]
    const int lift_off = fold ? P.off.lift : P.off.phi_face;

    // ---- off-tile neighbour traces: one (off-tile face, component) item per
    // lane for the first XMAX off-tile faces (the flux items of any further
    // ones form their neighbour's trace themselves).  The neighbour columns
    // are requested before the rows are written out of their registers;
    // the own traces are formed while they arrive.
    const int n_gh = n_off < XMAX ? n_off : XMAX;
    const bool gl = t < 4 * n_gh;
    const int gk = t >> 2, gc = t & 3, gsl = gl ? sExt[gk] : 0;
    const int gkey = gl ? (sIn[gsl] & 3) : 0;  // the neighbour's local face
    double xn[N];
    {
        const double* ur = Uin + (gl ? static_cast<int64_t>(sNb[gsl]) * NS : 0) + gc;
#pragma unroll
        for (int j = 0; j < NI; ++j) xn[j] = ur[nbr_node<C>(gkey, j) * 4];
    }
    L::put_rows(sW, t, ru);
    wave_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u[i] = sW[el * L::SIO + i * 4 + s];
        acc[i] = 0.0;
    }
    wave_sync();  // the staging rows are reused below
    __builtin_amdgcn_s_setprio(0);
    if (t < TE) sRd[t] = 1.0 / sGe[t * 5 + 4];
    // ---- own traces on every face, canonical point order
    {
        const double (*cols[1])[N] = {&u};
#pragma unroll 1
        for (int f = 0; f < NF; ++f)
            trace_face_rt<C, 1>(f, T + P.off.phi_face, cols, [&](int q, const double* a) {
                sT[atq(f * NQF + q, el) + s] = a[0];
            });
    }
    if (gl) {
        const double* ph = T + P.off.phi_face + gkey * N * NQF;
        const double (*cn[1])[N] = {&xn};
        trace_nbr<C, 1>(gkey, [&](int k) { return ph[k]; }, cn,
                        [&](int q, const double* a) { sG[(gk * NQF + q) * 4 + gc] = a[0]; });
    }
    wave_sync();
    // ---- one Riemann solve per edge point.  The rounds are wave-uniform so
    // that every lane takes part in the cross-lane reads of the edge
    // geometry (a lane outside the execution mask would return zero).
    const int n_it = n_items * NQF;
#pragma unroll 1
    for (int base = 0; base < n_it; base += DG_BS) {
        const int it = base + t;
        const bool act = it < n_it;
        const int ui = act ? it / NQF : 0, q = act ? it - ui * NQF : 0;
        const int A = sItem[ui];
        const double e_nx = __shfl(g_nx, A, 64), e_ny = __shfl(g_ny, A, 64);
        const double e_jac = __shfl(g_jac, A, 64);
        if (!act) continue;
        const int elA = A / NF, fA = A - elA * NF;
        const int infA = sIn[A], sdA = (infA >> 2) & 1, pl = sPlan[A];
        const int64_t k = infA >> 3;
        const int qA = sdA ? NQF - 1 - q : q;
        double* pa = sT + atq(fA * NQF + qA, elA);
        double* pb = nullptr;
        double fx[4], sp;
        if (sNb[A] < 0) {  // boundary: weighted flux precomputed by k_bflux
            const double* bf = P.bflux + ((k - P.n_int) * NQF + q) * 5;
#pragma unroll
            for (int c = 0; c < 4; ++c) fx[c] = bf[c];
            sp = bf[4];
        } else {
            // The other side's trace: the partner's row (a partner in the
            // patch is side 1 of the owner's edge), a ghost row, or, past the
            // ghost rows, the same sums from global memory.  The two sides
            // are read through swapped LDS addresses rather than swapped
            // values.
            const int kx = -2 - pl, qn = NQF - 1 - qA;
            const double* po = nullptr;
            if (pl >= 0) {
                const int elB = pl / NF, fB = pl - elB * NF;
                po = pb = sT + atq(fB * NQF + (NQF - 1 - q), elB);
            } else if (kx < XMAX) {
                po = sG + (kx * NQF + qn) * 4;
            }
            double ul[4], ur[4];
            if (po) {
                const double* pl_ = sdA == 0 ? pa : po;
                const double* pr_ = sdA == 0 ? po : pa;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    ul[c] = pl_[c];
                    ur[c] = pr_[c];
                }
            } else {
                const double* un = Uin + static_cast<int64_t>(sNb[A]) * NS;
                const double* ph = T + P.off.phi_face + (infA & 3) * N * NQF + qn;
#pragma unroll 1
                for (int c = 0; c < 4; ++c) {
                    double v = 0.0;
                    for (int j = 0; j < NI; ++j) {
                        const int i = nbr_node<C>(infA & 3, j);
                        v += ph[i * NQF] * un[i * 4 + c];
                    }
                    const double a = pa[c];
                    ul[c] = sdA == 0 ? a : v;
                    ur[c] = sdA == 0 ? v : a;
                }
            }
            sp = interface_flux<FLUX>(P, ul, ur, e_nx, e_ny, fx);
            const double sc = e_jac * sWq[q];
#pragma unroll
            for (int c = 0; c < 4; ++c) fx[c] = fx[c] * sc;
        }
        const unsigned long long spb = __double_as_longlong(sp > 0.0 ? sp : 0.0);
        double sg = sdA ? -1.0 : 1.0;
        if (fold) sg = sdA ? -sRd[elA] : sRd[elA];  // = sg / det J exactly
#pragma unroll
        for (int c = 0; c < 4; ++c) pa[c] = sg * fx[c];
        atomicMax(&sSp[A], spb);
        double len = 0.0;  // the edge's length weight, for both sides' wave-speed sums
        if (q == 0) {
            for (int qq = 0; qq < NQF; ++qq) len += e_jac * T[P.off.w_face + qq];
            sLen[A] = len;
        }
        if (pb) {
            const int elB = pl / NF;
            const double sgB = fold ? -sRd[elB] : -1.0;
#pragma unroll
            for (int c = 0; c < 4; ++c) pb[c] = sgB * fx[c];
            atomicMax(&sSp[pl], spb);
            if (q == 0) sLen[pl] = len;
        }
    }
    wave_sync();
    // ---- lift, face by face, and the wave-speed sum
    double wave = 0.0;
#pragma unroll 1
    for (int f = 0; f < NF; ++f) {
        double fw[NQF];
#pragma unroll
        for (int q = 0; q < NQF; ++q) fw[q] = sT[atq(f * NQF + q, el) + s];
        const double* lb = T + lift_off + f * N * NQF;
#pragma unroll
        for (int i = 0; i < N; ++i) {
#pragma unroll
            for (int q = 0; q < NQF; ++q) acc[i] += lb[i * NQF + q] * fw[q];
        }
        if (live) {
            const int sl = el * NF + f;
            wave += __longlong_as_double(static_cast<long long>(sSp[sl])) * sLen[sl];
        }
    }
    wave_sync();

    dg_volume<C, L>(P, T, sW, sGe, u, acc, t, cs, ne, fold,
                    [&](int pe) -> int64_t { return sId[pe]; });
    dg_epilogue<C, L, STAGE>(
        P, T, sW, sIn, u, acc, wave, t, el, s, e, live, ne, slot, dt, dt_factor, ca, cb, mode,
        partial, flags, counters,
        [&]() {  // the staged rows back, as they were loaded
            const int tq = opaque(t);
#pragma unroll
            for (int j = 0; j < L::NR; ++j) {
                const int x = tq + j * DG_BS, r = x / L::EP;
                if (r < ne)
                    __builtin_nontemporal_store(
                        reinterpret_cast<const d2*>(sW)[x],
                        reinterpret_cast<d2*>(Uout + static_cast<int64_t>(sId[r]) * NS) +
                            (x - r * L::EP));
            }
        },
        [&](typename L::Row (&rn)[L::NR]) {  // the rows of Un, as U was loaded
            const int tq = opaque(t);
#pragma unroll
            for (int j = 0; j < L::NR; ++j) {
                const int x = tq + j * DG_BS, r = x / L::EP;
                const d2* src = reinterpret_cast<const d2*>(
                    Un + static_cast<int64_t>(sId[r < TE ? r : 0]) * NS);
                rn[j] = r < ne ? src[x - r * L::EP] : d2{0.0, 0.0};
            }
        });
}

// ------------------------------------------------- boundary-face pre-pass
// One item per (boundary edge, face point): owner trace, boundary flux
// (inflow / outflow / wall / axis), weighted by |dX/dsigma| w_q, plus the
// signal speed.  Runs before k_dg on the same stream; it keeps the boundary
// physics out of the main kernel's register budget.
template <class C>
__global__ __launch_bounds__(BS) void k_bflux(const nz_problem P, const double* __restrict__ T,
                                              const double* __restrict__ U) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int64_t nb = P.n_edge - P.n_int;
    if (it >= nb * C::NQF) return;
    const int64_t b = it / C::NQF;
    const int q = static_cast<int>(it - b * C::NQF);
    const int64_t e = P.bedge[2 * b];
    const int ft = P.bedge[2 * b + 1];
    const int f = ft >> 3, tag = ft & 7;
    double uo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < C::N; ++i) {
        const double w = T[P.off.phi_face + (f * C::N + i) * C::NQF + q];
        for (int s = 0; s < 4; ++s) uo[s] += w * U[(e * C::N + i) * 4 + s];
    }
    const int64_t k = P.n_int + b;
    const double* eg = C::STR ? P.edge_geo + k * 3 : P.edge_geo + (k * C::NQF + q) * 3;
    double fx[4];
    const double nx = eg[0], ny = eg[1], pb = P.p_back;
    const double sp = boundary_flux(P, tag, uo, nx, ny, pb, fx);
    const double sc = eg[2] * T[P.off.w_face + q];
    double* out = P.bflux + (b * C::NQF + q) * 5;
    for (int s = 0; s < 4; ++s) out[s] = fx[s] * sc;
    out[4] = sp;
}

// ------------------------------------------------------- standalone limiter
template <class C>
__global__ __launch_bounds__(BS) void k_limit(const nz_problem P, const double* __restrict__ T,
                                              double* __restrict__ U, uint8_t* __restrict__ flags,
                                              unsigned long long* __restrict__ counters) {
    __shared__ double sPF[C::PF];
    for (int it = threadIdx.x; it < C::PF; it += BS) sPF[it] = T[P.off.phi_face + it];
    __syncthreads();
    const int64_t e = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (e >= P.n_elem) return;
    uint8_t fl = 0;
    if (P.limiter >= 1 && C::N > 1) {
        int side[C::NF];
        for (int f = 0; f < C::NF; ++f) side[f] = (P.fe_info[e * C::NF + f] >> 2) & 1;
        double mw[C::N];
        for (int i = 0; i < C::N; ++i) mw[i] = C::AFF ? T[P.off.mean_w + i] : P.mean_w[e * C::N + i];
        const LimTables tb{T + P.off.phi_vol, sPF};
        int scaled = 0;
        fl = limit_element<C>(U + e * C::NS, 1, mw, tb, side, P, &scaled);
        if (counters) {
            wave_count(&counters[0], scaled != 0);
            wave_count(&counters[1], (fl & 1) != 0);
        }
    }
    if (flags) flags[e] = fl;
}

// ------------------------------------------------------------ superbee
// Superbee TVB slope limiter (src/spatial/limiter.py:246-405) in three
// passes: cell means and per-block maxima of |mean_s|; the global maxima;
// then per element the trace-mean increments against the neighbours' mean
// jumps, one theta per element, applied in place.
__device__ __forceinline__ double minmod(double a, double b) {
    return a * b <= 0.0 ? 0.0 : (fabs(a) < fabs(b) ? a : b);
}
__device__ __forceinline__ double maxmod(double a, double b) {
    return a * b <= 0.0 ? 0.0 : (fabs(a) > fabs(b) ? a : b);
}

template <class C>
__global__ __launch_bounds__(BS) void k_sb_means(const nz_problem P, const double* __restrict__ T,
                                                 const double* __restrict__ U,
                                                 double* __restrict__ ubar,
                                                 double* __restrict__ pmax) {
    __shared__ double red[4][BS / 64];
    const int64_t e = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    if (e < P.n_elem) {
        double ub[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < C::N; ++i) {
            const double w = C::AFF ? T[P.off.mean_w + i] : P.mean_w[e * C::N + i];
            for (int s = 0; s < 4; ++s) ub[s] += w * U[(e * C::N + i) * 4 + s];
        }
        for (int s = 0; s < 4; ++s) {
            ubar[e * 4 + s] = ub[s];
            m[s] = fabs(ub[s]);
        }
    }
    for (int s = 0; s < 4; ++s) {
        double v = m[s];
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) red[s][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int w = 0; w < BS / 64; ++w) v = fmax(v, red[threadIdx.x][w]);
        pmax[blockIdx.x * 4 + threadIdx.x] = v;
    }
}

__global__ __launch_bounds__(BS) void k_sb_scale(const double* __restrict__ pmax, int64_t nblk,
                                                 double* __restrict__ scale) {
    __shared__ double red[4][BS / 64];
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t b = threadIdx.x; b < nblk; b += BS)
        for (int s = 0; s < 4; ++s) m[s] = fmax(m[s], pmax[b * 4 + s]);
    for (int s = 0; s < 4; ++s) {
        double v = m[s];
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) red[s][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int w = 0; w < BS / 64; ++w) v = fmax(v, red[threadIdx.x][w]);
        scale[threadIdx.x] = v;
    }
}

template <class C>
__global__ __launch_bounds__(BS) void k_sb_apply(const nz_problem P, const double* __restrict__ T,
                                                 double* __restrict__ U,
                                                 const double* __restrict__ ubar,
                                                 const double* __restrict__ scale, double tvb,
                                                 double area_total,
                                                 unsigned long long* __restrict__ counters) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (e >= P.n_elem) return;
    constexpr double tol = 1e-12;
    double wsum = 0.0;
    for (int q = 0; q < C::NQF; ++q) wsum += T[P.off.w_face + q];
    const double share = P.area[e] / area_total;
    double ub[4];
    for (int s = 0; s < 4; ++s) ub[s] = ubar[e * 4 + s];
    double theta = 1.0;
    for (int f = 0; f < C::NF; ++f) {
        const int nb = P.fe_nbr[e * C::NF + f];
        const int side = (P.fe_info[e * C::NF + f] >> 2) & 1;
        const double* ph = T + P.off.phi_face + (side * C::NF + f) * C::N * C::NQF;
        for (int s = 0; s < 4; ++s) {
            double tm = 0.0;
            for (int i = 0; i < C::N; ++i) {
                const double u = U[(e * C::N + i) * 4 + s];
                for (int q = 0; q < C::NQF; ++q) tm += ph[i * C::NQF + q] * T[P.off.w_face + q] * u;
            }
            const double dev = tm / wsum - ub[s];
            if (nb < 0 || fabs(dev) <= tol) continue;  // boundary face or no increment
            if (fabs(dev) <= tvb * share * scale[s]) continue;  // below the TVB threshold
            const double jump = ubar[static_cast<int64_t>(nb) * 4 + s] - ub[s];
            const double allowed = maxmod(minmod(2.0 * jump, dev), minmod(jump, 2.0 * dev));
            const double tf = fmin(fmax(allowed / dev, 0.0), 1.0);
            theta = tf < theta ? tf : theta;
        }
    }
    if (theta < 1.0) {
        for (int i = 0; i < C::N; ++i)
            for (int s = 0; s < 4; ++s) {
                double* x = &U[(e * C::N + i) * 4 + s];
                *x = ub[s] + theta * (*x - ub[s]);
            }
    }
    if (counters) wave_count(&counters[0], theta < 1.0);
}

// ---------------------------------------------------------------- adjoint
// Vector-Jacobian product of the unfolded residual, lambda^T dR/dU, and the
// cotangents of the geometry it depends on (per-point metrics, edge normals
// and Jacobians, back pressure).  Flux, boundary and Euler-flux Jacobians are
// recomputed per point with forward-mode duals and contracted at once; no
// Jacobian is stored.  Structure as k_dg: a tile of TE elements per
// workgroup, the element's own U and lambda staged in LDS, the neighbours'
// traces gathered from HBM, register accumulators per (element, basis).
//
// With R_i = -sum_q wdet_q [D^xi_iq (xi_x F + xi_y G) + D^eta_iq (eta_x F + eta_y G)]
//            + sum_f sign sum_q phi_f,iq Fhat_q |dX/dsigma|_q w_q
// (src/spatial/assembly.py:52-157), the VJP is
//   volume: g_q = -(b^x_q . dF/dU + b^y_q . dG/dU), b^x = wdet (a^xi xi_x + a^eta eta_x),
//           a^xi_s = sum_i lambda_is D^xi_iq; dU_j += sum_q phi_jq g_q;
//   faces:  a = |dX/dsigma| w (Lambda_owner - Lambda_neighbour) at each point,
//           dU_own,j += phi_own,jq a . dFhat/dU_own.

enum AdjMode { kAdjVjp = 0, kAdjRk = 1, kAdjDefect = 2 };

// Boundary faces: a . dF_b/dU_own per point, plus edge-geometry and
// back-pressure cotangents.  One item per (boundary edge, point).
template <class C>
__global__ __launch_bounds__(BS) void k_bvjp(const nz_problem P, const double* __restrict__ T,
                                             const double* __restrict__ U,
                                             const double* __restrict__ L,
                                             double* __restrict__ bv,
                                             double* __restrict__ gedge,
                                             double* __restrict__ gpb) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int64_t nb = P.n_edge - P.n_int;
    if (it >= nb * C::NQF) return;
    const int64_t b = it / C::NQF;
    const int q = static_cast<int>(it - b * C::NQF);
    const int64_t e = P.bedge[2 * b];
    const int ft = P.bedge[2 * b + 1];
    const int f = ft >> 3, tag = ft & 7;
    double uo[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < C::N; ++i) {
        const double w = T[P.off.phi_face + (f * C::N + i) * C::NQF + q];
        for (int s = 0; s < 4; ++s) {
            uo[s] += w * U[(e * C::N + i) * 4 + s];
            lo[s] += w * L[(e * C::N + i) * 4 + s];
        }
    }
    const int64_t k = P.n_int + b;
    const double* eg = C::STR ? P.edge_geo + k * 3 : P.edge_geo + (k * C::NQF + q) * 3;
    using D = Dual<7>;
    D ud[4];
    for (int s = 0; s < 4; ++s) ud[s] = D::seed(uo[s], s);
    const D nx = D::seed(eg[0], 4), ny = D::seed(eg[1], 5), pb = D::seed(P.p_back, 6);
    D fx[4];
    boundary_flux(P, tag, ud, nx, ny, pb, fx);
    const double wq = T[P.off.w_face + q];
    const double jw = eg[2] * wq;
    double v[7] = {0, 0, 0, 0, 0, 0, 0}, gj = 0.0;
    for (int s = 0; s < 4; ++s) {
        const double a = jw * lo[s];
        for (int d = 0; d < 7; ++d) v[d] += a * fx[s].d[d];
        gj += lo[s] * fx[s].v;
    }
    for (int d = 0; d < 4; ++d) bv[(b * C::NQF + q) * 4 + d] = v[d];
    if (gedge) {
        double* ge = gedge + (k * C::NQF + q) * 3;
        ge[0] = v[4];
        ge[1] = v[5];
        ge[2] = gj * wq;
    }
    if (gpb) gpb[b * C::NQF + q] = v[6];
}

// The boundary-flux Jacobians of U for a march of cotangents at fixed U
// (nz_problem.bjac): per boundary face point, J[s][d] = dFb_s / du_d at the
// owner's trace (forward-mode duals, as k_bvjp differentiates the flux) and
// the point's |dX/dsigma| w_q.  The one-wave VJP kernel applies it to the
// cotangent's traces, so the march launches no boundary pre-pass.
template <class C>
__global__ __launch_bounds__(BS) void k_bjac(const nz_problem P, const double* __restrict__ T,
                                             const double* __restrict__ U,
                                             double* __restrict__ bjac) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int64_t nb = P.n_edge - P.n_int;
    if (it >= nb * C::NQF) return;
    const int64_t b = it / C::NQF;
    const int q = static_cast<int>(it - b * C::NQF);
    const int64_t e = P.bedge[2 * b];
    const int ft = P.bedge[2 * b + 1];
    const int f = ft >> 3, tag = ft & 7;
    double uo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < C::N; ++i) {
        const double w = T[P.off.phi_face + (f * C::N + i) * C::NQF + q];
        for (int s = 0; s < 4; ++s) uo[s] += w * U[(e * C::N + i) * 4 + s];
    }
    const int64_t k = P.n_int + b;
    const double* eg = C::STR ? P.edge_geo + k * 3 : P.edge_geo + (k * C::NQF + q) * 3;
    using D = Dual<4>;
    D ud[4];
    for (int s = 0; s < 4; ++s) ud[s] = D::seed(uo[s], s);
    const D nx(eg[0]), ny(eg[1]), pb(P.p_back);
    D fx[4];
    boundary_flux(P, tag, ud, nx, ny, pb, fx);
    double* out = bjac + (b * C::NQF + q) * BJAC_DOUBLES;
    for (int s = 0; s < 4; ++s)
        for (int d = 0; d < 4; ++d) out[s * 4 + d] = fx[s].d[d];
    out[16] = eg[2] * T[P.off.w_face + q];
    out[17] = out[18] = out[19] = 0.0;
}

// The register budget of the reverse-mode Roe instantiation is held to the
// occupancy the LDS allows (three waves per SIMD at P2: no spills, and 19 %
// faster than the unconstrained allocation's two); the dual-number
// instantiations need far more registers and are left unconstrained.
template <class C, int FLUX, bool GEO>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(FLUX == 0 && !GEO ? C::AWPE : 1)))
void k_vjp(const nz_problem P, const double* __restrict__ T,
                                            const double* __restrict__ U,
                                            const double* __restrict__ Lin,
                                            const double* __restrict__ bv,
                                            double* __restrict__ out, double* __restrict__ gmet,
                                            double* __restrict__ gedge, int mode,
                                            const double* __restrict__ seed,
                                            const double* __restrict__ Lbase,
                                            const double* __restrict__ dt,
                                            double* __restrict__ acc_rk, double* __restrict__ Lnext,
                                            int rk_stage) {
    constexpr int N = C::N, NQV = C::NQV, NQF = C::NQF, NF = C::NF, NS = C::NS;
    constexpr int TE = C::TE, TP = C::TP, SLOTS = C::SLOTS;
    // tangent directions per dual pass: all of them (the owner's state, and
    // the normal when geometry cotangents are asked for) in one pass
    constexpr int KB = GEO ? 6 : 4;
    __shared__ double sU[NS * TP];
    __shared__ double sL[NS * TP];
    __shared__ double sW[C::AWORK];
    __shared__ double sPF[C::PF];
    __shared__ double sG[5 * TE];
    __shared__ int sSide[NF * TE];

    const int t = threadIdx.x;
    const int el = t % TE;
    const int slot = wave_slot() / (TE / 64);
    int64_t e0;
    int ne;
    if (P.vjp_tiles) {  // a subset of the mesh: the live designs of a batch
        e0 = P.vjp_tiles[2 * blockIdx.x];
        ne = static_cast<int>(P.vjp_tiles[2 * blockIdx.x + 1] - e0);
    } else {
        e0 = static_cast<int64_t>(blockIdx.x) * TE;
        ne = static_cast<int>(P.n_elem - e0 < TE ? P.n_elem - e0 : TE);
    }
    const bool live = el < ne;
    const int64_t e = e0 + el;

    for (int it = t; it < TE * NS; it += BS) {
        const int r = it / NS, c = it - r * NS;
        sU[c * TP + r] = r < ne ? U[(e0 + r) * NS + c] : 0.0;
        sL[c * TP + r] = r < ne ? Lin[(e0 + r) * NS + c] : 0.0;
    }
    for (int it = t; it < C::PF; it += BS) sPF[it] = T[P.off.phi_face + it];
    if (t < TE && live) {
        if (C::AFF)
            for (int k = 0; k < 5; ++k) sG[k * TE + t] = P.geo[e * 5 + k];
        for (int f = 0; f < NF; ++f) sSide[f * TE + t] = (P.fe_info[e * NF + f] >> 2) & 1;
    }
    __syncthreads();

    double acc[C::NIT][4];
#pragma unroll
    for (int k = 0; k < C::NIT; ++k)
        for (int s = 0; s < 4; ++s) acc[k][s] = 0.0;

    // ---------------------------------------------------------------- faces
    double* sTn = sW;                     // [NQF][4][TP] neighbour U trace
    double* sLn = sW + NQF * 4 * TP;      // [NQF][4][TP] neighbour lambda trace
    double* sV = sW + 2 * NQF * 4 * TP;   // [NQF][4][TP] a . dFhat/dU_own
#pragma unroll 1
    for (int f = 0; f < NF; ++f) {
#pragma unroll 1
        for (int s = slot; s < 4; s += SLOTS) {
            double tu[NQF], tl[NQF];
#pragma unroll
            for (int q = 0; q < NQF; ++q) tu[q] = tl[q] = 0.0;
            if (live) {
                const int nb = P.fe_nbr[e * NF + f];
                if (nb >= 0) {
                    const int info = P.fe_info[e * NF + f];
                    const int of = info & 3, oside = 1 - ((info >> 2) & 1);
                    const double* ur = U + static_cast<int64_t>(nb) * NS + s;
                    const double* lr = Lin + static_cast<int64_t>(nb) * NS + s;
                    const double* ph = sPF + (oside * NF + of) * N * NQF;
                    for (int j = 0; j < (C::SPT ? C::FN : N); ++j) {
                        const int i = nbr_node<C>(of, j);
                        const double x = ur[i * 4], y = lr[i * 4];
#pragma unroll
                        for (int q = 0; q < NQF; ++q) {
                            tu[q] += ph[i * NQF + q] * x;
                            tl[q] += ph[i * NQF + q] * y;
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NQF; ++q) {
                sTn[(q * 4 + s) * TP + el] = tu[q];
                sLn[(q * 4 + s) * TP + el] = tl[q];
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int q = slot; q < NQF; q += SLOTS) {
            if (live) {
                const int nb = P.fe_nbr[e * NF + f];
                const int info = P.fe_info[e * NF + f];
                const int side = (info >> 2) & 1;
                const int64_t k = info >> 3;
                double v[4];
                if (nb >= 0) {
                    const double* ph = sPF + (side * NF + f) * N * NQF + q;
                    double uo[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
                    for (int j = 0; j < (C::SPT ? C::FN : N); ++j) {
                        const int i = nbr_node<C>(f, j);
                        const double b = ph[i * NQF];
                        for (int s = 0; s < 4; ++s) {
                            uo[s] += b * sU[(i * 4 + s) * TP + el];
                            lo[s] += b * sL[(i * 4 + s) * TP + el];
                        }
                    }
                    const double* eg = C::STR ? P.edge_geo + k * 3 : P.edge_geo + (k * NQF + q) * 3;
                    const double wq = T[P.off.w_face + q];
                    const double jw = eg[2] * wq;
                    // Tangents: the owner's state, and the normal when its
                    // cotangent is asked for, KB directions per pass (each
                    // pass re-evaluates the flux with a KB-wide dual).
                    constexpr int KD = GEO ? 6 : 4;
                    using D = Dual<KB>;
                    double dl[4], un[4];  // dl = Lambda_owner - Lambda_neighbour
                    for (int s = 0; s < 4; ++s) {
                        un[s] = sTn[(q * 4 + s) * TP + el];
                        const double ln = sLn[(q * 4 + s) * TP + el];
                        dl[s] = side == 0 ? lo[s] - ln : ln - lo[s];
                    }
                    double g[KD], gj = 0.0;
                    if constexpr (FLUX == 0 && !GEO) {
                        // Roe: one reverse-mode pass for the owner's cotangent
                        double uL[4], uR[4], ob[4];
                        for (int s = 0; s < 4; ++s) {
                            uL[s] = side == 0 ? uo[s] : un[s];
                            uR[s] = side == 0 ? un[s] : uo[s];
                            ob[s] = jw * dl[s];
                        }
                        roe_vjp(uL, uR, eg[0], eg[1], P.gamma, P.efix, side == 0, ob, g);
                    } else {
#pragma unroll
                    for (int d0 = 0; d0 < KD; d0 += KB) {
                        D ul[4], ur[4], fx[4];
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            const D own = seeded<KB>(uo[s], s - d0), nbr(un[s]);
                            ul[s] = side == 0 ? own : nbr;
                            ur[s] = side == 0 ? nbr : own;
                        }
                        if constexpr (GEO) {
                            const D nx = seeded<KB>(eg[0], 4 - d0), ny = seeded<KB>(eg[1], 5 - d0);
                            interface_flux<FLUX>(P, ul, ur, nx, ny, fx);
                        } else {
                            interface_flux<FLUX>(P, ul, ur, eg[0], eg[1], fx);
                        }
#pragma unroll
                        for (int j = 0; j < KB; ++j) {
                            double a = 0.0;
                            for (int s = 0; s < 4; ++s) a += (jw * dl[s]) * fx[s].d[j];
                            g[d0 + j] = a;
                        }
                        if (d0 == 0)
                            for (int s = 0; s < 4; ++s) gj += dl[s] * fx[s].v;
                    }
                    }
                    for (int d = 0; d < 4; ++d) v[d] = g[d];
                    if constexpr (GEO) {
                        if (gedge && side == 0) {
                            double* ge = gedge + (k * NQF + q) * 3;
                            ge[0] = g[4];
                            ge[1] = g[5];
                            ge[2] = gj * wq;
                        }
                    }
                } else {
                    const double* b = bv + ((k - P.n_int) * NQF + q) * 4;
                    for (int d = 0; d < 4; ++d) v[d] = b[d];
                }
                for (int d = 0; d < 4; ++d) sV[(q * 4 + d) * TP + el] = v[d];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < C::NIT; ++kk) {
            const int i = slot + kk * SLOTS;
            if (i < N && live && on_face<C>(f, i)) {
                const int side = sSide[f * TE + el];
                for (int q = 0; q < NQF; ++q) {
                    const double b0 = T[P.off.phi_face + ((0 * NF + f) * N + i) * NQF + q];
                    const double b1 = T[P.off.phi_face + ((1 * NF + f) * N + i) * NQF + q];
                    const double b = side ? b1 : b0;
                    for (int d = 0; d < 4; ++d) acc[kk][d] += b * sV[(q * 4 + d) * TP + el];
                }
            }
        }
        __syncthreads();
    }

    // --------------------------------------------------------------- volume
    double* sGq = sW;  // [QR][4][TP] cotangent on the point state
#pragma unroll 1
    for (int q0 = 0; q0 < NQV; q0 += C::QR) {
        {
            const int qq = slot, q = q0 + qq;
            if (q < NQV && live) {
                double uq[4] = {0, 0, 0, 0}, ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
                for (int i = 0; i < N; ++i) {
                    const double b = T[P.off.phi_vol + i * NQV + q];
                    const double dx = T[P.off.dxi + i * NQV + q], dy = T[P.off.deta + i * NQV + q];
                    for (int s = 0; s < 4; ++s) {
                        uq[s] += b * sU[(i * 4 + s) * TP + el];
                        const double l = sL[(i * 4 + s) * TP + el];
                        ax[s] += l * dx;
                        ay[s] += l * dy;
                    }
                }
                double xx, xy, ex, ey, wd;
                if (C::AFF) {
                    xx = sG[el];
                    xy = sG[TE + el];
                    ex = sG[2 * TE + el];
                    ey = sG[3 * TE + el];
                    wd = T[P.off.w_vol + q] * sG[4 * TE + el];
                } else {
                    const double* gq = P.geo + (e * NQV + q) * 5;
                    xx = gq[0];
                    xy = gq[1];
                    ex = gq[2];
                    ey = gq[3];
                    wd = gq[4];
                }
                double g[4] = {0, 0, 0, 0};
                double gm[5] = {0, 0, 0, 0, 0};
                double bx[4], by[4];
                for (int s = 0; s < 4; ++s) {
                    bx[s] = wd * (ax[s] * xx + ay[s] * ex);
                    by[s] = wd * (ax[s] * xy + ay[s] * ey);
                }
                double Fv[4], Gv[4];
                if constexpr (GEO) {
                    // the fluxes' values are needed for the metric cotangents
                    using D = Dual<4>;
                    D ud[4];
                    for (int s = 0; s < 4; ++s) ud[s] = D::seed(uq[s], s);
                    const Prim<D> w = primitives(ud, P.gamma);
                    const D F[4] = {w.r * w.u, w.r * w.u * w.u + w.p, w.r * w.u * w.v, w.r * w.u * w.H};
                    const D G[4] = {w.r * w.v, w.r * w.u * w.v, w.r * w.v * w.v + w.p, w.r * w.v * w.H};
                    for (int s = 0; s < 4; ++s) {
                        for (int d = 0; d < 4; ++d) g[d] -= bx[s] * F[s].d[d] + by[s] * G[s].d[d];
                        Fv[s] = F[s].v;
                        Gv[s] = G[s].v;
                    }
                } else {  // reverse mode: one pass instead of four tangents
                    double ub[4];
                    euler_flux_vjp(uq, P.gamma, bx, by, ub);
                    for (int d = 0; d < 4; ++d) g[d] = -ub[d];
                }
                for (int s = 0; s < 4; ++s) {
                    if constexpr (GEO) {
                        gm[0] -= wd * ax[s] * Fv[s];
                        gm[1] -= wd * ax[s] * Gv[s];
                        gm[2] -= wd * ay[s] * Fv[s];
                        gm[3] -= wd * ay[s] * Gv[s];
                        gm[4] -= ax[s] * (xx * Fv[s] + xy * Gv[s]) + ay[s] * (ex * Fv[s] + ey * Gv[s]);
                    }
                }
                for (int d = 0; d < 4; ++d) sGq[(qq * 4 + d) * TP + el] = g[d];
                if (GEO && gmet)
                    for (int d = 0; d < 5; ++d) gmet[(e * NQV + q) * 5 + d] = gm[d];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < C::NIT; ++kk) {
            const int i = slot + kk * SLOTS;
            if (i < N && live) {
                for (int qq = 0; qq < C::QR && q0 + qq < NQV; ++qq) {
                    const double b = T[P.off.phi_vol + i * NQV + q0 + qq];
                    for (int d = 0; d < 4; ++d) acc[kk][d] += b * sGq[(qq * 4 + d) * TP + el];
                }
            }
        }
        __syncthreads();
    }

    if (mode == kAdjVjp) {
#pragma unroll
        for (int kk = 0; kk < C::NIT; ++kk) {
            const int i = slot + kk * SLOTS;
            if (i < N && live)
                for (int s = 0; s < 4; ++s) out[(e * N + i) * 4 + s] = acc[kk][s];
        }
        return;
    }
    if (mode == kAdjDefect) {  // partial sums of (vjp - seed)^2; `out` holds one per block
        __shared__ double sRed[BS / 64];
        double sq = 0.0;
#pragma unroll
        for (int kk = 0; kk < C::NIT; ++kk) {
            const int i = slot + kk * SLOTS;
            if (i < N && live)
                for (int s = 0; s < 4; ++s) {
                    const double d = acc[kk][s] - seed[(e * N + i) * 4 + s];
                    sq += d * d;
                }
        }
        const double tot = block_sum(sq, sRed);
        if (t == 0) out[blockIdx.x] = tot;
        return;
    }

    // adjoint pseudo-time rate -M^-1 (vjp - seed) and one classical RK4 stage
    // (src/sensitivity.py:344-357)
    double* sR = sW;
#pragma unroll
    for (int kk = 0; kk < C::NIT; ++kk) {
        const int i = slot + kk * SLOTS;
        if (i < N)
            for (int s = 0; s < 4; ++s)
                sR[(i * 4 + s) * TP + el] = live ? acc[kk][s] - seed[(e * N + i) * 4 + s] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < C::NIT; ++kk) {
        const int i = slot + kk * SLOTS;
        if (i < N && live) {
            double r[4] = {0, 0, 0, 0};
            const double rdet = C::AFF ? 1.0 / sG[4 * TE + el] : 0.0;  // M^-1 = M_ref^-1 / det J
            for (int j = 0; j < N; ++j) {
                const double m = C::AFF ? T[P.off.minv_ref + i * N + j] * rdet
                                        : P.minv[e * sym_size(N) + sym_index(N, i, j)];
                for (int s = 0; s < 4; ++s) r[s] += m * sR[(j * 4 + s) * TP + el];
            }
            const double h = dt[e];
            for (int s = 0; s < 4; ++s) {
                const int64_t x = (e * N + i) * 4 + s;
                const double kv = -r[s];
                const double lb = Lbase[x];
                if (rk_stage == 1) {
                    acc_rk[x] = kv;
                    Lnext[x] = lb + 0.5 * h * kv;
                } else if (rk_stage == 2) {
                    acc_rk[x] = acc_rk[x] + 2.0 * kv;
                    Lnext[x] = lb + 0.5 * h * kv;
                } else if (rk_stage == 3) {
                    acc_rk[x] = acc_rk[x] + 2.0 * kv;
                    Lnext[x] = lb + h * kv;
                } else {
                    Lnext[x] = lb + h / 6.0 * (acc_rk[x] + kv);
                }
            }
        }
    }
}

// ------------------------------------------- residual VJP on wave tiles
// (dR/dU)^T lambda, state cotangents only, on the mapping of k_dg: one wave
// per 16-element tile, four lanes per element, each lane holding its
// element's columns U[:, s] and lambda[:, s] and its accumulator in registers
// and reading reference tables with scalar loads.  Per face the lanes form
// the own and neighbour traces of U and lambda in canonical point order; one
// (point, element) item per lane pulls jw (lambda_own - lambda_nbr) back
// through the interface flux with a hand-written reverse-mode pass
// (roe_vjp, hllc_vjp, ausm_vjp), and the element lifts the result with the
// side-0 face table.  Boundary faces take the pre-pass
// values (k_bvjp).  Per volume round the lanes interpolate U and the
// reference gradients of lambda to four points; one item per lane pulls the
// contravariant cotangents back through the Euler fluxes (euler_flux_vjp) and
// the element projects them with phi.  Modes and outputs as k_vjp.
template <class C>
struct VTile {
    static constexpr int N = C::N, NQV = C::NQV, NQF = C::NQF, NF = C::NF, NS = C::NS;
    static constexpr int QR = DG_BS / TE;
    static constexpr int FACE = 4 * NQF * QS;  // own U, own lambda, neighbour U, neighbour lambda
    static constexpr int VOL = 3 * QR * QS;    // U, dlambda/dxi, dlambda/deta at the round's points
    static constexpr int IO = 2 * TE * NS;     // the U and lambda tiles as in HBM
    static constexpr int WORK = FACE > VOL ? (FACE > IO ? FACE : IO) : (VOL > IO ? VOL : IO);
    static constexpr int GF = C::STR ? 3 : 3 * NQF;
    static constexpr int EGN = TE * NF * GF;
    static constexpr int GEN = C::AFF ? TE * 5 : 1;
    // register budget: three waves per SIMD (168 VGPRs) up to P2 where the
    // LDS allows them (measured 2 % faster than two waves at P2), two above
    static constexpr int LDS = 8 * (WORK + NF * N * NQF + EGN + GEN + TE) + 8 * TE * NF;
    static constexpr int WPE = N <= 3 && LDS * 12 <= 160 * 1024 ? 3 : 2;
    // HLLC's reverse pass is capped at two waves per SIMD (256 registers): at
    // three it spills and runs 60 % slower
};

template <class C, int FLUX>
__global__ __launch_bounds__(DG_BS) __attribute__((amdgpu_waves_per_eu(FLUX == 1 ? 2 : VTile<C>::WPE)))
void k_vjpw(const nz_problem P, const double* __restrict__ T,
                                              const double* __restrict__ U,
                                              const double* __restrict__ Lin,
                                              const double* __restrict__ bv,
                                              double* __restrict__ out, int mode,
                                              const double* __restrict__ seed,
                                              const double* __restrict__ Lbase,
                                              const double* __restrict__ dt,
                                              double* __restrict__ acc_rk,
                                              double* __restrict__ Lnext, int rk_stage) {
    using L = VTile<C>;
    constexpr int N = C::N, NQV = C::NQV, NQF = C::NQF, NF = C::NF, NS = C::NS;
    constexpr int NPF = NF * N * NQF;
    constexpr int NU = TE * NS / DG_BS;  // = N
    constexpr int NI = C::SPT ? C::FN : N;  // neighbour coefficients a face trace reads
    __shared__ double sW[L::WORK];
    __shared__ double sPF[NPF];
    __shared__ double sEG[L::EGN];
    __shared__ double sGe[L::GEN];
    __shared__ double sDt[TE];
    __shared__ double sWv[NQV];  // volume quadrature weights
    __shared__ int sNb[TE * NF], sIn[TE * NF];

    const int t = threadIdx.x;
    const int el = t >> 2, s = t & 3;
    static_assert(NQV <= DG_BS, "one lane per volume weight");
    if (t < NQV) sWv[t] = T[P.off.w_vol + t];
    // The tile: an entry of the adjoint's own tile table (a batch's live
    // designs), else of the element kernel's locality table (the neighbour
    // gathers of a spatially coherent launch hit L2; the slot keeps the
    // defect's partial sums in element order), else consecutive.
    int64_t e0, slot;
    int ne;
    if (P.vjp_tiles) {
        e0 = P.vjp_tiles[2 * blockIdx.x];
        ne = static_cast<int>(P.vjp_tiles[2 * blockIdx.x + 1] - e0);
        slot = blockIdx.x;
    } else if (P.tiles && P.tiles_cover_mesh) {
        e0 = P.tiles[3 * blockIdx.x];
        ne = static_cast<int>(P.tiles[3 * blockIdx.x + 1] - e0);
        slot = P.tiles[3 * blockIdx.x + 2];
    } else {
        e0 = static_cast<int64_t>(blockIdx.x) * TE;
        ne = static_cast<int>(P.n_elem - e0 < TE ? P.n_elem - e0 : TE);
        slot = blockIdx.x;
    }
    const bool live = el < ne;
    const int64_t e = e0 + el;
    const int cs = at(0, s, el);

    __builtin_amdgcn_s_setprio(2);  // staging at raised priority, as in k_dg
    // ---- stage U and lambda (coalesced), the maps, the face table, the metrics
    {
        // unconditional loads (masked past the tile) and the edge geometry in
        // the same batch, from the face map in lane r, as in k_dg
        static_assert(TE * NF <= DG_BS, "one lane per element face");
        constexpr int NT = (NPF + DG_BS - 1) / DG_BS;
        constexpr int NG = (L::GEN + DG_BS - 1) / DG_BS;
        constexpr int NE_ = (L::EGN + DG_BS - 1) / DG_BS;
        double ru[NU], rl[NU], rt[NT], rg[NG], rq[NE_];
        int rn, ri;
        {
            const bool ok = t < ne * NF;
            const int a = P.fe_nbr[e0 * NF + (ok ? t : 0)], b = P.fe_info[e0 * NF + (ok ? t : 0)];
            rn = ok ? a : -1;
            ri = ok ? b : 0;
        }
#pragma unroll
        for (int j = 0; j < NE_; ++j) {
            const int it = t + j * DG_BS;
            const int r = it / L::GF, jj = it - r * L::GF;
            const bool ok = it < L::EGN && r < ne * NF;
            const int64_t k = __shfl(ri, ok ? r : 0, 64) >> 3;
            const double v = P.edge_geo[(ok ? k : 0) * L::GF + (ok ? jj : 0)];
            rq[j] = ok ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int it = t + j * DG_BS, r = it / NS;
            const double a = U[e0 * NS + (r < ne ? it : 0)];
            const double b = Lin[e0 * NS + (r < ne ? it : 0)];
            ru[j] = r < ne ? a : 0.0;
            rl[j] = r < ne ? b : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int it = t + j * DG_BS;
            const double v = T[P.off.phi_face + (it < NPF ? it : 0)];
            rt[j] = it < NPF ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            const int it = t + j * DG_BS;
            const bool ok = C::AFF && it < ne * 5;
            const double v = C::AFF ? P.geo[e0 * 5 + (ok ? it : 0)] : 0.0;
            rg[j] = ok ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            sW[t + j * DG_BS] = ru[j];
            sW[TE * NS + t + j * DG_BS] = rl[j];
        }
        if (t < TE * NF) {
            sNb[t] = rn;
            sIn[t] = ri;
        }
#pragma unroll
        for (int j = 0; j < NT; ++j)
            if (t + j * DG_BS < NPF) sPF[t + j * DG_BS] = rt[j];
#pragma unroll
        for (int j = 0; j < NG; ++j)
            if (t + j * DG_BS < L::GEN) sGe[t + j * DG_BS] = rg[j];
#pragma unroll
        for (int j = 0; j < NE_; ++j)
            if (t + j * DG_BS < L::EGN) sEG[t + j * DG_BS] = rq[j];
    }
    wave_sync();
    __builtin_amdgcn_s_setprio(0);
    double u[N], l[N], acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u[i] = sW[el * NS + i * 4 + s];
        l[i] = sW[TE * NS + el * NS + i * 4 + s];
        acc[i] = 0.0;
    }
    wave_sync();

    double* sUo = sW;                 // own U trace, then the pulled-back cotangent
    double* sLo = sW + NQF * QS;      // own lambda trace
    double* sUn = sW + 2 * NQF * QS;  // neighbour U trace
    double* sLn = sW + 3 * NQF * QS;  // neighbour lambda trace

    // ---------------------------------------------------------------- faces
#pragma unroll 1
    for (int f = 0; f < NF; ++f) {
        const int nb = sNb[el * NF + f], inf = sIn[el * NF + f];
        const bool has = live && nb >= 0;
        const int key = has ? (inf & 3) : -1;  // the neighbour's local face
        double xn[N], yn[N];
        {
            const int64_t base = (has ? static_cast<int64_t>(nb) : e0) * NS + s;
            const int kk = has ? key : 0;
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const int i = nbr_node<C>(kk, j);
                xn[j] = U[base + i * 4];
                yn[j] = Lin[base + i * 4];
            }
        }
        {
            const double (*cols[2])[N] = {&u, &l};
            trace_face_rt<C, 2>(f, T + P.off.phi_face, cols, [&](int q, const double* a) {
                sUo[q * QS + cs] = a[0];
                sLo[q * QS + cs] = a[1];
            });
        }
        {
            double tu[NQF], tl[NQF];
#pragma unroll
            for (int q = 0; q < NQF; ++q) tu[q] = tl[q] = 0.0;
            const int key0 = __builtin_amdgcn_readfirstlane(key);
            const double (*cn[2])[N] = {&xn, &yn};
            auto put = [&](int q, const double* a) {
                if constexpr (C::ROLL_T) {
                    sUn[q * QS + cs] = a[0];
                    sLn[q * QS + cs] = a[1];
                } else {
                    tu[q] = a[0];
                    tl[q] = a[1];
                }
            };
            if (__all(key == key0) && key0 >= 0) {
                const double* ph = T + P.off.phi_face + key0 * N * NQF;
                trace_nbr<C, 2>(key0, [&](int k) { return ph[k]; }, cn, put);
            } else if (has) {
                const double* ph = sPF + key * N * NQF;
                trace_nbr<C, 2>(key, [&](int k) { return ph[k]; }, cn, put);
            } else if constexpr (C::ROLL_T) {
#pragma unroll
                for (int q = 0; q < NQF; ++q) sUn[q * QS + cs] = sLn[q * QS + cs] = 0.0;
            }
            if constexpr (!C::ROLL_T) {
#pragma unroll
                for (int q = 0; q < NQF; ++q) {
                    sUn[q * QS + cs] = tu[q];
                    sLn[q * QS + cs] = tl[q];
                }
            }
        }
        wave_sync();
        // (q,e): the cotangent at edge point q, one item per lane
        for (int it = t; it < NQF * TE; it += DG_BS) {
            const int q = it / TE, pe = it % TE;
            if (pe < ne) {
                const int nbp = sNb[pe * NF + f], infp = sIn[pe * NF + f];
                const int sd = (infp >> 2) & 1;
                const int64_t k = infp >> 3;
                const int qo = sd ? NQF - 1 - q : q;  // this element's canonical index
                double v[4];
                if (nbp >= 0) {
                    const int qn = NQF - 1 - qo;
                    const double* eg = sEG + (pe * NF + f) * L::GF + (C::STR ? 0 : q * 3);
                    const double jw = eg[2] * face_weight<C>(T, P, q);
                    double uL[4], uR[4], ob[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const double uo = sUo[at(qo, c, pe)], un = sUn[at(qn, c, pe)];
                        const double lo = sLo[at(qo, c, pe)], ln = sLn[at(qn, c, pe)];
                        uL[c] = sd == 0 ? uo : un;
                        uR[c] = sd == 0 ? un : uo;
                        ob[c] = jw * (sd == 0 ? lo - ln : ln - lo);
                    }
                    if constexpr (FLUX == 0) {
                        roe_vjp(uL, uR, eg[0], eg[1], P.gamma, P.efix, sd == 0, ob, v);
                    } else if constexpr (FLUX == 2) {
                        ausm_vjp(uL, uR, eg[0], eg[1], P.gamma, sd == 0, ob, v);
                    } else {
                        hllc_vjp(uL, uR, eg[0], eg[1], P.gamma, P.low_mach, sd == 0, ob, v);
                    }
                } else if (P.bjac) {  // boundary faces at a fixed U: J^T (jw lambda)
                    const double* jb = P.bjac + ((k - P.n_int) * NQF + q) * BJAC_DOUBLES;
                    const double jw = jb[16];
#pragma unroll
                    for (int d = 0; d < 4; ++d) v[d] = 0.0;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const double a = jw * sLo[at(qo, c, pe)];
#pragma unroll
                        for (int d = 0; d < 4; ++d) v[d] += a * jb[c * 4 + d];
                    }
                } else {  // boundary faces: the pre-pass's values
                    const double* b = bv + ((k - P.n_int) * NQF + q) * 4;
#pragma unroll
                    for (int c = 0; c < 4; ++c) v[c] = b[c];
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) sUo[at(qo, c, pe)] = v[c];
            }
        }
        wave_sync();
        if constexpr (C::ROLL_T) {
            lift_face_rolled<C>(T + P.off.phi_face + f * N * NQF,
                                [&](int q) { return sUo[q * QS + cs]; }, acc);
        } else {
            double fw[NQF];
#pragma unroll
            for (int q = 0; q < NQF; ++q) fw[q] = sUo[q * QS + cs];
            lift_face_rt<C>(f, T + P.off.phi_face, fw, acc);
        }
        // the next face rewrites these rows after the reads above
    }
    wave_sync();

    // --------------------------------------------------------------- volume
    {
        double* sQu = sW;                   // U at the round's points, then the cotangent
        double* sQx = sW + L::QR * QS;      // dlambda/dxi
        double* sQy = sW + 2 * L::QR * QS;  // dlambda/deta
        const double* vt = T + P.off.vol_t;  // per point: phi, dxi, deta, ... (N each)
#pragma unroll 1
        for (int q0 = 0; q0 < NQV; q0 += L::QR) {
            // per-point metrics of this lane's item, requested before the
            // interpolation (unconditional, at a valid point)
            double gm[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if constexpr (!C::AFF) {
                const int qq = t / TE, pe = t % TE, q = q0 + qq;
                const bool ok = q < NQV && pe < ne;
                const double* gq = P.geo + ((e0 + (ok ? pe : 0)) * NQV + (ok ? q : 0)) * 5;
#pragma unroll
                for (int k = 0; k < 5; ++k) gm[k] = gq[k];
            }
#pragma unroll 1
            for (int qq = 0; qq < L::QR; ++qq) {
                if (q0 + qq < NQV) {
                    const double* ph = vt + (q0 + qq) * 5 * N;
                    double a = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        a += ph[i] * u[i];
                        bx += ph[N + i] * l[i];
                        by += ph[2 * N + i] * l[i];
                    }
                    sQu[qq * QS + cs] = a;
                    sQx[qq * QS + cs] = bx;
                    sQy[qq * QS + cs] = by;
                }
            }
            wave_sync();
            {
                const int qq = t / TE, pe = t % TE, q = q0 + qq;
                if (q < NQV && pe < ne) {
                    double uq[4], fb[4], gb[4], ub[4];
                    double xx, xy, ex, ey, wd;
                    if (C::AFF) {
                        const double* ge = sGe + pe * 5;
                        xx = ge[0];
                        xy = ge[1];
                        ex = ge[2];
                        ey = ge[3];
                        wd = sWv[q] * ge[4];
                    } else {
                        xx = gm[0];
                        xy = gm[1];
                        ex = gm[2];
                        ey = gm[3];
                        wd = gm[4];
                    }
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        uq[c] = sQu[at(qq, c, pe)];
                        const double ax = sQx[at(qq, c, pe)], ay = sQy[at(qq, c, pe)];
                        fb[c] = wd * (ax * xx + ay * ex);
                        gb[c] = wd * (ax * xy + ay * ey);
                    }
                    euler_flux_vjp(uq, P.gamma, fb, gb, ub);
#pragma unroll
                    for (int c = 0; c < 4; ++c) sQu[at(qq, c, pe)] = -ub[c];
                }
            }
            wave_sync();
#pragma unroll 1
            for (int qq = 0; qq < L::QR; ++qq) {
                const int q = q0 + qq;
                if (q < NQV) {
                    const double a = sQu[qq * QS + cs];
                    const double* ph = vt + q * 5 * N;
#pragma unroll
                    for (int i = 0; i < N; ++i) acc[i] += ph[i] * a;
                }
            }
            // the next round rewrites these rows after the reads above
        }
    }
    wave_sync();

    // ------------------------------------------------------------ epilogue
    if (mode == kAdjVjp) {
#pragma unroll
        for (int i = 0; i < N; ++i) sW[el * NS + i * 4 + s] = acc[i];
        wave_sync();
        if (ne == TE) {
#pragma unroll
            for (int j = 0; j < NU; ++j) out[e0 * NS + t + j * DG_BS] = sW[t + j * DG_BS];
        } else {
            for (int it = t; it < ne * NS; it += DG_BS) out[e0 * NS + it] = sW[it];
        }
        return;
    }
    {  // the seed tile, coalesced
        double rs[NU];
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int it = t + j * DG_BS;
            rs[j] = it / NS < ne ? seed[e0 * NS + it] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) sW[TE * NS + t + j * DG_BS] = rs[j];
        if (s == 0) sDt[el] = live && mode == kAdjRk ? dt[e] : 0.0;
    }
    wave_sync();
    double d[N];
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = acc[i] - sW[TE * NS + el * NS + i * 4 + s];
    if (mode == kAdjDefect) {  // partial sums of (vjp - seed)^2, one per tile
        double sq = 0.0;
        if (live)
            for (int i = 0; i < N; ++i) sq += d[i] * d[i];
        for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
        if (t == 0) out[slot] = sq;
        return;
    }
    // adjoint pseudo-time rate -M^-1 (vjp - seed) and one classical RK4
    // stage (src/sensitivity.py:344-357), the update in coalesced rows.
    // Affine elements: M^-1 = M_ref^-1 / det J, applied as M_ref^-1 times
    // 1 / det J (one division per element instead of one per entry).
    const double rdet = C::AFF ? 1.0 / sGe[el * 5 + 4] : 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double r = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double m = C::AFF ? T[P.off.minv_ref + i * N + j] * rdet
                                    : (live ? P.minv[e * sym_size(N) + sym_index(N, i, j)] : 0.0);
            r += m * d[j];
        }
        sW[el * NS + i * 4 + s] = -r;
    }
    wave_sync();
    auto rk = [&](int64_t x, double kv, double h, double lb, double ac) {
        if (rk_stage == 1) {
            acc_rk[x] = kv;
            Lnext[x] = lb + 0.5 * h * kv;
        } else if (rk_stage == 2) {
            acc_rk[x] = ac + 2.0 * kv;
            Lnext[x] = lb + 0.5 * h * kv;
        } else if (rk_stage == 3) {
            acc_rk[x] = ac + 2.0 * kv;
            Lnext[x] = lb + h * kv;
        } else {
            Lnext[x] = lb + h / 6.0 * (ac + kv);
        }
    };
    if (ne == TE) {  // a full tile: every row requested before the first update
        double lb[NU], ac[NU];
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int64_t x = e0 * NS + t + j * DG_BS;
            lb[j] = Lbase[x];
            ac[j] = rk_stage == 1 ? 0.0 : acc_rk[x];
        }
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            const int it = t + j * DG_BS;
            rk(e0 * NS + it, sW[it], sDt[it / NS], lb[j], ac[j]);
        }
    } else {
        for (int it = t; it < ne * NS; it += DG_BS) {
            const int64_t x = e0 * NS + it;
            rk(x, sW[it], sDt[it / NS], Lbase[x], rk_stage == 1 ? 0.0 : acc_rk[x]);
        }
    }
}

// ------------------------------------------------------------ small kernels
template <class C>
__global__ __launch_bounds__(BS) void k_inverse_mass(const nz_problem P, const double* __restrict__ T,
                                                     const double* __restrict__ R,
                                                     double* __restrict__ out) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (it >= P.n_elem * C::N) return;
    const int64_t e = it / C::N;
    const int i = static_cast<int>(it - e * C::N);
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < C::N; ++j) {
        const double m = C::AFF ? T[P.off.minv_ref + i * C::N + j] / P.geo[e * 5 + 4]
                                : P.minv[e * sym_size(C::N) + sym_index(C::N, i, j)];
        for (int s = 0; s < 4; ++s) r[s] += m * R[(e * C::N + j) * 4 + s];
    }
    for (int s = 0; s < 4; ++s) out[it * 4 + s] = r[s];
}

// L2 projection of point values onto the basis (src/timestepping/
// initialize.py:41-46): U_e = M_e^-1 sum_q phi(q) (w det)_q v_q, one thread
// per (element, component) with the element's right-hand side in registers.
template <class C>
__global__ __launch_bounds__(BS) void k_project(const nz_problem P, const double* __restrict__ T,
                                                const double* __restrict__ V,
                                                double* __restrict__ out) {
    constexpr int N = C::N, NQV = C::NQV;
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (it >= P.n_elem * 4) return;
    const int64_t e = it >> 2;
    const int s = static_cast<int>(it & 3);
    double rhs[N];
#pragma unroll
    for (int i = 0; i < N; ++i) rhs[i] = 0.0;
    for (int q = 0; q < NQV; ++q) {
        const double wd = C::AFF ? T[P.off.w_vol + q] * P.geo[e * 5 + 4] : P.geo[(e * NQV + q) * 5 + 4];
        const double v = wd * V[(e * NQV + q) * 4 + s];
#pragma unroll
        for (int i = 0; i < N; ++i) rhs[i] += T[P.off.phi_vol + i * NQV + q] * v;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double r = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double m = C::AFF ? T[P.off.minv_ref + i * N + j] / P.geo[e * 5 + 4]
                                    : P.minv[e * sym_size(N) + sym_index(N, i, j)];
            r += m * rhs[j];
        }
        out[(e * N + i) * 4 + s] = r;
    }
}

// Boundary integrals of the owner-side trace for the observables
// (src/post/postprocess.py:56-75,155-171,193-226).  One workgroup scans every
// boundary point in a fixed order and reduces in LDS in a fixed tree, so the
// sums do not depend on the launch.  out[8]: axial momentum flux through the
// inflow and the outflow, -p n_x ds over the walls, mass flux through the
// inflow and the outflow, the outflow length, and the integrals of the Mach
// number and of p over the outflow.
template <class C>
__global__ __launch_bounds__(BS) void k_boundary_integrals(const nz_problem P,
                                                           const double* __restrict__ T,
                                                           const double* __restrict__ U,
                                                           double* __restrict__ out) {
    constexpr int N = C::N, NQF = C::NQF;
    constexpr int K = 8;
    __shared__ double red[K][BS];
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = 0.0;
    const int t = threadIdx.x;
    const int64_t nb = P.n_edge - P.n_int;
    for (int64_t it = t; it < nb * NQF; it += BS) {
        const int64_t b = it / NQF;
        const int q = static_cast<int>(it - b * NQF);
        const int64_t e = P.bedge[2 * b];
        const int ft = P.bedge[2 * b + 1];
        const int f = ft >> 3, tag = ft & 7;
        if (tag < 1 || tag > 3) continue;  // the symmetry axis carries none of them
        double S[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < N; ++i) {
            const double w = T[P.off.phi_face + (f * N + i) * NQF + q];
            for (int c = 0; c < 4; ++c) S[c] += w * U[(e * N + i) * 4 + c];
        }
        const int64_t k = P.n_int + b;
        const double* eg = P.edge_geo + (C::STR ? k * 3 : (k * NQF + q) * 3);
        const double nx = eg[0], ny = eg[1], ds = eg[2] * T[P.off.w_face + q];
        const double r = fmax(S[0], kFloor), vx = S[1] / r, vy = S[2] / r;
        const double p = fmax((P.gamma - 1.0) * (S[3] - 0.5 * r * (vx * vx + vy * vy)), kFloor);
        const double vn = vx * nx + vy * ny;
        if (tag == 3) {
            acc[2] += -p * nx * ds;
        } else {
            const int o = tag == 2 ? 1 : 0;
            acc[o] += (r * vx * vn + p * nx) * ds;
            acc[3 + o] += r * vn * ds;
            if (tag == 2) {
                acc[5] += ds;
                acc[6] += sqrt((vx * vx + vy * vy) * r / (P.gamma * p)) * ds;
                acc[7] += p * ds;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) red[j][t] = acc[j];
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int j = 0; j < K; ++j) red[j][t] += red[j][t + w];
        __syncthreads();
    }
    if (t < K) out[t] = red[t][0];
}

}  // namespace nz

// ---------------------------------------------------------------- generic kernels
namespace {

__global__ __launch_bounds__(nz::BS) void k_norm_partial(const double* __restrict__ A, int64_t n,
                                                         double w0, double w1, double w2, double w3,
                                                         double* __restrict__ partial) {
    __shared__ double red[nz::BS / 64];
    const double w[4] = {w0, w1, w2, w3};
    double acc = 0.0;
    for (int64_t it = static_cast<int64_t>(blockIdx.x) * nz::BS + threadIdx.x; it < n;
         it += static_cast<int64_t>(gridDim.x) * nz::BS) {
        const double v = A[it] * w[it & 3];
        acc += v * v;
    }
    const double tot = nz::block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// One workgroup sums the per-tile partials in a fixed order (four independent
// accumulators per thread keep loads in flight), so the norm is reproducible.
constexpr int FIN_BS = 1024;

__device__ void norm_of_partials(const double* __restrict__ partial, int64_t nparts, double count,
                                 double* __restrict__ out) {
    __shared__ double red[FIN_BS / 64];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t it = threadIdx.x;
    for (; it + 3 * FIN_BS < nparts; it += 4 * FIN_BS) {
        a0 += partial[it];
        a1 += partial[it + FIN_BS];
        a2 += partial[it + 2 * FIN_BS];
        a3 += partial[it + 3 * FIN_BS];
    }
    for (; it < nparts; it += FIN_BS) a0 += partial[it];
    double v = (a0 + a1) + (a2 + a3);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int t = threadIdx.x;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double tot = 0.0;
        for (int w = 0; w < FIN_BS / 64; ++w) tot += red[w];
        out[0] = sqrt(tot / count);
    }
}

__global__ __launch_bounds__(FIN_BS) void k_norm_finish(const double* __restrict__ partial,
                                                        int64_t nparts, double count,
                                                        double* __restrict__ out) {
    norm_of_partials(partial, nparts, count, out);
}

// one workgroup per segment of the partials
__global__ __launch_bounds__(FIN_BS) void k_norm_segments(const double* __restrict__ partial,
                                                          const int64_t* __restrict__ seg,
                                                          const double* __restrict__ counts,
                                                          double* __restrict__ out) {
    const int64_t d = blockIdx.x;
    norm_of_partials(partial + seg[d], seg[d + 1] - seg[d], counts[d], out + d);
}

__global__ __launch_bounds__(nz::BS) void k_time_step(const double* __restrict__ wave,
                                                      const double* __restrict__ area, int64_t n,
                                                      double factor, double* __restrict__ dt) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * nz::BS + threadIdx.x;
    if (e >= n) return;
    const double d = wave[e] < nz::kFloor ? nz::kFloor : wave[e];
    dt[e] = factor * 2.0 * area[e] / d;
}

thread_local char g_err[512] = "";

int fail(int code, const char* msg) {
    std::snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

int check_launch(const char* what) {
    const hipError_t st = hipGetLastError();
    if (st != hipSuccess) {
        std::snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(st));
        return NZ_E_LAUNCH;
    }
    return NZ_OK;
}

// ------------------------------------------------ static scratch (no allocation)
// The boundary pre-pass output of nz_rate_noalloc lives in this file-scope
// device array instead of a caller's buffer: up to kNoAllocDoubles doubles
// (8 MiB), i.e. (boundary edges) x (face points) <= 209,715.  Each GPU holds
// its own copy for the life of the process.  Host code never names the
// array in a kernel argument or copy: its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
namespace noalloc {
constexpr int64_t kNoAllocDoubles = int64_t{1} << 20;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_lock;
double* g_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, or nullptr with the
// error recorded
double* scratch_ptr() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fail(NZ_E_LAUNCH, "nz_rate_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> hold(g_lock);
    if (!g_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t st = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !ptr) {
            char msg[256];
            std::snprintf(msg, sizeof msg, "nz_rate_noalloc: hipGetSymbolAddress: %s",
                          hipGetErrorString(st));
            fail(NZ_E_LAUNCH, msg);
            return nullptr;
        }
        g_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_ptr[dev];
}
}  // namespace noalloc

// ------------------------------------------------------------- dispatch
struct Ops {
    void (*dg)(const nz_problem&, const double*, const double*, double*, double*, double, double,
               double, int, double*, uint8_t*, unsigned long long*, int64_t, hipStream_t);
    void (*limit)(const nz_problem&, double*, uint8_t*, unsigned long long*, hipStream_t);
    void (*inverse_mass)(const nz_problem&, const double*, double*, hipStream_t);
    void (*project)(const nz_problem&, const double*, double*, hipStream_t);
    void (*boundary_integrals)(const nz_problem&, const double*, double*, hipStream_t);
    void (*superbee)(const nz_problem&, double*, double, double, double*,
                     unsigned long long*, hipStream_t);
    void (*vjp)(const nz_problem&, const double*, const double*, double*, double*, double*,
                double*, double*, int, const double*, const double*, const double*, double*,
                double*, int, int64_t, hipStream_t);
    void (*bjac)(const nz_problem&, const double*, double*, hipStream_t);
    int te;
};

template <class C>
struct Launch {
    static void dg(const nz_problem& P, const double* Uin, const double* Un, double* Uout,
                   double* dt, double fac, double a, double b, int mode, double* partial,
                   uint8_t* flags, unsigned long long* counters, int64_t nblk, hipStream_t st) {
        (void)nblk;  // k_dg tiles are nz::TE elements, or come from the tile table
        const int64_t nblk_dg = P.tiles ? P.n_tiles : (P.n_elem + nz::TE - 1) / nz::TE;
        const int64_t nbq = (P.n_edge - P.n_int) * C::NQF;
        const bool stage = mode == nz::kStageFirst || mode == nz::kStageNext;
        // a chained stage finds Uin's boundary fluxes left by the one before
        if (nbq > 0 && !(stage && (P.bflux_chain & 1)))
            hipLaunchKernelGGL((nz::k_bflux<C>), dim3(static_cast<unsigned>((nbq + nz::BS - 1) / nz::BS)),
                               dim3(nz::BS), 0, st, P, P.tables, Uin);
        if constexpr (C::AFF) {
            if (P.patches) {
                if (P.n_patches <= 0) return;
                const dim3 gp(static_cast<unsigned>(P.n_patches));
#define NZ_DGP(FL, ST)                                                                             \
    hipLaunchKernelGGL((nz::k_dgp<C, FL, ST>), gp, dim3(nz::DG_BS), 0, st, P, P.tables, Uin, Un,    \
                       Uout, dt, fac, a, b, mode, partial, flags, counters)
                if (P.flux_id == 1) {
                    if (stage) NZ_DGP(1, true);
                    else NZ_DGP(1, false);
                } else if (P.flux_id == 2) {
                    if (stage) NZ_DGP(2, true);
                    else NZ_DGP(2, false);
                } else {
                    if (stage) NZ_DGP(0, true);
                    else NZ_DGP(0, false);
                }
#undef NZ_DGP
                return;
            }
        }
        if (nblk_dg <= 0) return;
        const dim3 g(static_cast<unsigned>(nblk_dg));
#define NZ_DG(FL, ST)                                                                              \
    hipLaunchKernelGGL((nz::k_dg<C, FL, ST>), g, dim3(nz::DG_BS), 0, st, P, P.tables, Uin, Un, Uout, \
                       dt, fac, a, b, mode, partial, flags, counters)
        if (P.flux_id == 1) {
            if (stage) NZ_DG(1, true);
            else NZ_DG(1, false);
        } else if (P.flux_id == 2) {
            if (stage) NZ_DG(2, true);
            else NZ_DG(2, false);
        } else {
            if (stage) NZ_DG(0, true);
            else NZ_DG(0, false);
        }
#undef NZ_DG
    }
    static void limit(const nz_problem& P, double* U, uint8_t* flags, unsigned long long* counters,
                      hipStream_t st) {
        const unsigned nb = static_cast<unsigned>((P.n_elem + nz::BS - 1) / nz::BS);
        hipLaunchKernelGGL((nz::k_limit<C>), dim3(nb), dim3(nz::BS), 0, st, P, P.tables, U, flags,
                           counters);
    }
    static void inverse_mass(const nz_problem& P, const double* R, double* out, hipStream_t st) {
        const unsigned nb = static_cast<unsigned>((P.n_elem * C::N + nz::BS - 1) / nz::BS);
        hipLaunchKernelGGL((nz::k_inverse_mass<C>), dim3(nb), dim3(nz::BS), 0, st, P, P.tables, R, out);
    }
    static void project(const nz_problem& P, const double* V, double* out, hipStream_t st) {
        const unsigned nb = static_cast<unsigned>((P.n_elem * 4 + nz::BS - 1) / nz::BS);
        hipLaunchKernelGGL((nz::k_project<C>), dim3(nb), dim3(nz::BS), 0, st, P, P.tables, V, out);
    }
    static void boundary_integrals(const nz_problem& P, const double* U, double* out,
                                   hipStream_t st) {
        hipLaunchKernelGGL((nz::k_boundary_integrals<C>), dim3(1), dim3(nz::BS), 0, st, P, P.tables,
                           U, out);
    }
    static void superbee(const nz_problem& P, double* U, double tvb, double area_total,
                         double* work, unsigned long long* counters, hipStream_t st) {
        const int64_t nb = (P.n_elem + nz::BS - 1) / nz::BS;
        double* ubar = work;
        double* pmax = work + P.n_elem * 4;
        double* scale = pmax + nb * 4;
        hipLaunchKernelGGL((nz::k_sb_means<C>), dim3(static_cast<unsigned>(nb)), dim3(nz::BS), 0, st,
                           P, P.tables, U, ubar, pmax);
        hipLaunchKernelGGL(nz::k_sb_scale, dim3(1), dim3(nz::BS), 0, st, pmax, nb, scale);
        hipLaunchKernelGGL((nz::k_sb_apply<C>), dim3(static_cast<unsigned>(nb)), dim3(nz::BS), 0, st,
                           P, P.tables, U, ubar, scale, tvb, area_total, counters);
    }
    static void vjp(const nz_problem& P, const double* U, const double* L, double* out,
                    double* gmet, double* gedge, double* gpb, double* bv, int mode,
                    const double* seed, const double* Lbase, const double* dt, double* acc,
                    double* Lnext, int stage, int64_t nblk, hipStream_t st) {
        const int64_t nbq = (P.n_edge - P.n_int) * C::NQF;
        const bool geo = gmet || gedge;  // geometry cotangents: the normal is differentiated too
        // the one-wave kernel applies a fixed U's boundary Jacobians itself
        const bool fixed = P.bjac && !geo && !gpb && P.vjp_kernel == 0;
        if (nbq > 0 && !fixed)
            hipLaunchKernelGGL((nz::k_bvjp<C>), dim3(static_cast<unsigned>((nbq + nz::BS - 1) / nz::BS)),
                               dim3(nz::BS), 0, st, P, P.tables, U, L, bv, gedge, gpb);
        if (nblk <= 0) return;  // an empty tile table: nothing to march
        if (!geo && P.vjp_kernel == 0) {  // state cotangents on wave tiles
#define NZ_VJPW(FL)                                                                               \
    hipLaunchKernelGGL((nz::k_vjpw<C, FL>), dim3(static_cast<unsigned>(nblk)), dim3(nz::DG_BS), 0, \
                       st, P, P.tables, U, L, bv, out, mode, seed, Lbase, dt, acc, Lnext, stage)
            if (P.flux_id == 1) NZ_VJPW(1);
            else if (P.flux_id == 2) NZ_VJPW(2);
            else NZ_VJPW(0);
#undef NZ_VJPW
            return;
        }
        // its own grid: entries of a tile table, or tiles of C::TE elements
        const dim3 g(static_cast<unsigned>(P.vjp_tiles ? P.n_vjp_tiles
                                                       : (P.n_elem + C::TE - 1) / C::TE));
#define NZ_VJP(FL)                                                                                \
    do {                                                                                          \
        if (geo)                                                                                  \
            hipLaunchKernelGGL((nz::k_vjp<C, FL, true>), g, dim3(nz::BS), 0, st, P, P.tables, U,  \
                               L, bv, out, gmet, gedge, mode, seed, Lbase, dt, acc, Lnext, stage); \
        else                                                                                      \
            hipLaunchKernelGGL((nz::k_vjp<C, FL, false>), g, dim3(nz::BS), 0, st, P, P.tables, U, \
                               L, bv, out, gmet, gedge, mode, seed, Lbase, dt, acc, Lnext, stage); \
    } while (0)
        if (P.flux_id == 1) NZ_VJP(1);
        else if (P.flux_id == 2) NZ_VJP(2);
        else NZ_VJP(0);
#undef NZ_VJP
    }
    static void bjac(const nz_problem& P, const double* U, double* out, hipStream_t st) {
        const int64_t nbq = (P.n_edge - P.n_int) * C::NQF;
        if (nbq > 0)
            hipLaunchKernelGGL((nz::k_bjac<C>), dim3(static_cast<unsigned>((nbq + nz::BS - 1) / nz::BS)),
                               dim3(nz::BS), 0, st, P, P.tables, U, out);
    }
    static Ops ops() {
        return Ops{&dg, &limit, &inverse_mass, &project, &boundary_integrals, &superbee, &vjp,
                   &bjac, C::TE};
    }
};

// (N, nqv, nqf, nface, affine, straight) for p = 0..3 of each element family
#define NZ_FAMILY(X)                                                              \
    X(1, 1, 1, 3, true, true) X(3, 4, 2, 3, true, true) X(6, 9, 3, 3, true, true) \
    X(10, 16, 4, 3, true, true)                                                   \
    X(1, 4, 2, 3, false, false) X(3, 9, 3, 3, false, false)                       \
    X(6, 16, 4, 3, false, false) X(10, 25, 5, 3, false, false)                    \
    X(1, 1, 1, 4, false, true) X(4, 4, 2, 4, false, true)                         \
    X(9, 9, 3, 4, false, true) X(16, 16, 4, 4, false, true)                       \
    X(1, 4, 2, 4, false, false) X(4, 9, 3, 4, false, false)                       \
    X(9, 16, 4, 4, false, false) X(16, 25, 5, 4, false, false)

bool lookup(const nz_problem* p, Ops* out) {
    if (!p) return false;
#define NZ_TRY(n, qv, qf, nf, aff, str)                                                  \
    if (p->nbf == n && p->nqv == qv && p->nqf == qf && p->nface == nf &&                 \
        (p->affine != 0) == aff && (p->straight != 0) == str) {                          \
        *out = Launch<nz::Cfg<n, qv, qf, nf, aff, str>>::ops();                          \
        return true;                                                                     \
    }
    NZ_FAMILY(NZ_TRY)
#undef NZ_TRY
    return false;
}

int resolve(const nz_problem* p, Ops* ops) {
    if (!p) return fail(NZ_E_ARG, "null problem");
    if (!lookup(p, ops)) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "unsupported element configuration: nface=%d N=%d nqv=%d nqf=%d affine=%d "
                      "straight=%d",
                      p->nface, p->nbf, p->nqv, p->nqf, p->affine, p->straight);
        return fail(NZ_E_ARG, msg);
    }
    if (p->n_elem <= 0) return fail(NZ_E_ARG, "empty mesh");
    if (p->flux_id < 0 || p->flux_id > 2) return fail(NZ_E_ARG, "flux_id must be 0, 1 or 2");
    if (p->n_int < 0 || p->n_int > p->n_edge) return fail(NZ_E_ARG, "n_int out of range");
    if (p->n_edge > p->n_int && (!p->bedge || !p->bflux))
        return fail(NZ_E_ARG, "boundary edges need bedge and the bflux workspace");
    if (p->tiles && (p->n_tiles < 0 || p->n_tiles > (int64_t{1} << 31)))
        return fail(NZ_E_ARG, "n_tiles out of range");
    if (p->tiles_cover_mesh != 0 && p->tiles_cover_mesh != 1)
        return fail(NZ_E_ARG, "tiles_cover_mesh must be 0 or 1");
    if (p->bflux_chain < 0 || p->bflux_chain > 3)
        return fail(NZ_E_ARG, "bflux_chain must be 0..3");
    if (p->bflux_chain != 0 && p->limiter == 2)
        return fail(NZ_E_ARG, "bflux_chain needs the stage's output final (limiter 2 runs passes after it)");
    if (p->tiles && p->tiles_cover_mesh && p->n_tiles != (p->n_elem + nz::TE - 1) / nz::TE)
        return fail(NZ_E_ARG, "tiles_cover_mesh needs one tile-table entry per tile of the mesh");
    if (p->vjp_tiles && (p->n_vjp_tiles < 0 || p->n_vjp_tiles > (int64_t{1} << 31)))
        return fail(NZ_E_ARG, "n_vjp_tiles out of range");
    if (p->vjp_kernel != 0 && p->vjp_kernel != 1) return fail(NZ_E_ARG, "vjp_kernel must be 0 or 1");
    if (p->patches && (p->n_patches < 0 || p->n_patches > (int64_t{1} << 31)))
        return fail(NZ_E_ARG, "n_patches out of range");
    return NZ_OK;
}

int64_t n_blocks(const nz_problem* p, const Ops& o) { return (p->n_elem + o.te - 1) / o.te; }
// elements per workgroup of the state-cotangent VJP launches: the wave-tile
// kernel's 16, or the four-wave kernel's tile
int64_t vjp_te(const nz_problem* p, const Ops& o) { return p->vjp_kernel == 0 ? nz::TE : o.te; }
int64_t n_blocks_vjp(const nz_problem* p, const Ops& o) {
    return p->vjp_tiles ? p->n_vjp_tiles : (p->n_elem + vjp_te(p, o) - 1) / vjp_te(p, o);
}
// norm partials of the element kernel: one per patch, tile-table entry or tile
int64_t n_blocks_dg(const nz_problem* p) {
    if (p->patches && p->affine) return p->n_patches;
    return p->tiles ? p->n_tiles : (p->n_elem + nz::TE - 1) / nz::TE;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* nz_last_error(void) { return g_err; }

int64_t nz_scratch_doubles(const nz_problem* p) {
    Ops o;
    if (resolve(p, &o) != NZ_OK) return -1;
    int64_t a = n_blocks(p, o) > n_blocks_dg(p) ? n_blocks(p, o) : n_blocks_dg(p);
    if (n_blocks_vjp(p, o) > a) a = n_blocks_vjp(p, o);
    return a > 1024 ? a : 1024;
}

int nz_face_node(int nface, int p, int f, int j) {
    if ((nface != 3 && nface != 4) || p < 1 || p > 3 || f < 0 || f >= nface || j < 0 || j > p)
        return -1;
    return nz::face_node(nface, p, f, j);
}

int64_t nz_adjoint_tile_elements(const nz_problem* p) {
    Ops o;
    if (resolve(p, &o) != NZ_OK) return -1;
    return vjp_te(p, o);
}

int nz_residual(const nz_problem* p, const double* U, double* R, double* wave, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (p->bflux_chain != 0) return fail(NZ_E_ARG, "bflux_chain is for nz_stage only");
    if (!U || !R || !wave) return fail(NZ_E_ARG, "null array");
    o.dg(*p, U, nullptr, R, wave, 0.0, 0.0, 0.0, nz::kRes, nullptr, nullptr, nullptr, n_blocks(p, o),
         static_cast<hipStream_t>(stream));
    return check_launch("nz_residual");
}

int nz_rate(const nz_problem* p, const double* U, double* F, double* wave, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (p->bflux_chain != 0) return fail(NZ_E_ARG, "bflux_chain is for nz_stage only");
    if (!U || !F || !wave) return fail(NZ_E_ARG, "null array");
    o.dg(*p, U, nullptr, F, wave, 0.0, 0.0, 0.0, nz::kRate, nullptr, nullptr, nullptr,
         n_blocks(p, o), static_cast<hipStream_t>(stream));
    return check_launch("nz_rate");
}

int64_t nz_rate_noalloc_max_boundary_points(void) { return noalloc::kNoAllocDoubles / 5; }

int nz_rate_noalloc(const nz_problem* p, const double* U, double* F, double* wave, void* stream) {
    if (!p) return fail(NZ_E_ARG, "null problem");
    if (p->bflux_chain != 0) return fail(NZ_E_ARG, "bflux_chain is for nz_stage only");
    if (!U || !F || !wave) return fail(NZ_E_ARG, "null array");
    const int64_t nbq = (p->n_edge - p->n_int) * static_cast<int64_t>(p->nqf);
    if (p->n_edge < p->n_int || p->nqf <= 0) return fail(NZ_E_ARG, "n_int out of range");
    if (nbq * 5 > noalloc::kNoAllocDoubles) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "nz_rate_noalloc: %lld boundary face points need %lld doubles of scratch; "
                      "the static scratch holds %lld (use nz_rate with a bflux workspace)",
                      static_cast<long long>(nbq), static_cast<long long>(nbq * 5),
                      static_cast<long long>(noalloc::kNoAllocDoubles));
        return fail(NZ_E_ARG, msg);
    }
    double* ws = noalloc::scratch_ptr();
    if (!ws) return NZ_E_LAUNCH;
    nz_problem q = *p;
    q.bflux = ws;  // every entry the element kernel reads is written first by k_bflux
    Ops o;
    int rc = resolve(&q, &o);
    if (rc) return rc;
    o.dg(q, U, nullptr, F, wave, 0.0, 0.0, 0.0, nz::kRate, nullptr, nullptr, nullptr,
         n_blocks(&q, o), static_cast<hipStream_t>(stream));
    return check_launch("nz_rate_noalloc");
}

int nz_inverse_mass(const nz_problem* p, const double* R, double* out, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!R || !out) return fail(NZ_E_ARG, "null array");
    o.inverse_mass(*p, R, out, static_cast<hipStream_t>(stream));
    return check_launch("nz_inverse_mass");
}

int nz_project(const nz_problem* p, const double* values, double* U, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!values || !U) return fail(NZ_E_ARG, "null array");
    o.project(*p, values, U, static_cast<hipStream_t>(stream));
    return check_launch("nz_project");
}

int nz_boundary_integrals(const nz_problem* p, const double* U, double* out, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U || !out) return fail(NZ_E_ARG, "null array");
    o.boundary_integrals(*p, U, out, static_cast<hipStream_t>(stream));
    return check_launch("nz_boundary_integrals");
}

int nz_stage(const nz_problem* p, const double* Uin, const double* Un, double* Uout, double* dt,
             double dt_factor, double a, double b, int first, double* scratch, uint8_t* flags,
             uint64_t* counters, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!Uin || !Uout || !dt) return fail(NZ_E_ARG, "null array");
    if (Uout == Uin || (Un && Uout == Un)) return fail(NZ_E_ARG, "Uout must not alias Uin or Un");
    if (!first && !Un) return fail(NZ_E_ARG, "Un is required after the first stage");
    if (first && !scratch) return fail(NZ_E_ARG, "scratch is required for the first stage");
    o.dg(*p, Uin, Un, Uout, dt, dt_factor, a, b, first ? nz::kStageFirst : nz::kStageNext, scratch,
         flags, reinterpret_cast<unsigned long long*>(counters), n_blocks(p, o),
         static_cast<hipStream_t>(stream));
    return check_launch("nz_stage");
}

int nz_norm_finish(const nz_problem* p, const double* scratch, double* out, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!scratch || !out) return fail(NZ_E_ARG, "null array");
    const double count = static_cast<double>(p->n_elem) * p->nbf * 4.0;
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(FIN_BS), 0, static_cast<hipStream_t>(stream),
                       scratch, n_blocks_dg(p), count, out);
    return check_launch("nz_norm_finish");
}

int nz_norm_segments(const nz_problem* p, const double* scratch, const int64_t* seg,
                     const double* counts, int64_t n_seg, double* out, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (n_seg < 0) return fail(NZ_E_ARG, "n_seg must be non-negative");
    if (n_seg == 0) return NZ_OK;
    if (!scratch || !seg || !counts || !out) return fail(NZ_E_ARG, "null array");
    hipLaunchKernelGGL(k_norm_segments, dim3(static_cast<unsigned>(n_seg)), dim3(FIN_BS), 0,
                       static_cast<hipStream_t>(stream), scratch, seg, counts, out);
    return check_launch("nz_norm_segments");
}

int nz_norm(const nz_problem* p, const double* A, double* scratch, double* out, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!A || !scratch || !out) return fail(NZ_E_ARG, "null array");
    const int64_t n = p->n_elem * p->nbf * 4;
    int64_t nb = (n + nz::BS - 1) / nz::BS;
    if (nb > 1024) nb = 1024;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_norm_partial, dim3(static_cast<unsigned>(nb)), dim3(nz::BS), 0, st, A, n,
                       p->wnorm[0], p->wnorm[1], p->wnorm[2], p->wnorm[3], scratch);
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(FIN_BS), 0, st, scratch, nb,
                       static_cast<double>(n), out);
    return check_launch("nz_norm");
}

int nz_limit(const nz_problem* p, double* U, uint8_t* flags, uint64_t* counters, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U) return fail(NZ_E_ARG, "null array");
    o.limit(*p, U, flags, reinterpret_cast<unsigned long long*>(counters),
            static_cast<hipStream_t>(stream));
    return check_launch("nz_limit");
}

int64_t nz_adjoint_work_doubles(const nz_problem* p) {
    if (!p || p->n_elem <= 0) return fail(NZ_E_ARG, "empty mesh");
    const int64_t nb = p->n_edge - p->n_int;
    return (nb > 0 ? nb : 1) * p->nqf * 4;
}

int nz_residual_vjp(const nz_problem* p, const double* U, const double* lam, double* dU,
                    double* gmet, double* gedge, double* gpb, double* work, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U || !lam || !dU || !work) return fail(NZ_E_ARG, "null array");
    o.vjp(*p, U, lam, dU, gmet, gedge, gpb, work, nz::kAdjVjp, nullptr, nullptr, nullptr, nullptr,
          nullptr, 0, n_blocks_vjp(p, o), static_cast<hipStream_t>(stream));
    return check_launch("nz_residual_vjp");
}

int nz_adjoint_rk4_stage(const nz_problem* p, const double* U, const double* lam_in,
                         const double* lam_base, const double* seed, const double* dt,
                         double* acc, double* lam_next, int stage, double* work, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U || !lam_in || !lam_base || !seed || !dt || !acc || !lam_next || !work)
        return fail(NZ_E_ARG, "null array");
    if (stage < 1 || stage > 4) return fail(NZ_E_ARG, "stage must be 1..4");
    if (lam_next == lam_in || lam_next == lam_base)
        return fail(NZ_E_ARG, "lam_next must not alias lam_in or lam_base");
    o.vjp(*p, U, lam_in, nullptr, nullptr, nullptr, nullptr, work, nz::kAdjRk, seed, lam_base, dt,
          acc, lam_next, stage, n_blocks_vjp(p, o), static_cast<hipStream_t>(stream));
    return check_launch("nz_adjoint_rk4_stage");
}

int nz_adjoint_defect(const nz_problem* p, const double* U, const double* lam, const double* seed,
                      double* work, double* scratch, double* out, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U || !lam || !seed || !work || !scratch || !out) return fail(NZ_E_ARG, "null array");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nb = n_blocks_vjp(p, o);
    o.vjp(*p, U, lam, scratch, nullptr, nullptr, nullptr, work, nz::kAdjDefect, seed, nullptr,
          nullptr, nullptr, nullptr, 0, nb, st);
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(FIN_BS), 0, st, scratch, nb,
                       static_cast<double>(p->n_elem) * p->nbf * 4.0, out);
    return check_launch("nz_adjoint_defect");
}

int nz_boundary_jacobian(const nz_problem* p, const double* U, double* bjac, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U || !bjac) return fail(NZ_E_ARG, "null array");
    o.bjac(*p, U, bjac, static_cast<hipStream_t>(stream));
    return check_launch("nz_boundary_jacobian");
}

int64_t nz_superbee_work_doubles(const nz_problem* p) {
    if (!p || p->n_elem <= 0) return fail(NZ_E_ARG, "empty mesh");
    return p->n_elem * 4 + ((p->n_elem + nz::BS - 1) / nz::BS) * 4 + 4;
}

int nz_superbee(const nz_problem* p, double* U, double tvb, double area_total, double* work,
                uint64_t* counters, void* stream) {
    Ops o;
    int rc = resolve(p, &o);
    if (rc) return rc;
    if (!U || !work) return fail(NZ_E_ARG, "null array");
    if (!(area_total > 0.0)) return fail(NZ_E_ARG, "area_total must be positive");
    o.superbee(*p, U, tvb, area_total, work, reinterpret_cast<unsigned long long*>(counters),
               static_cast<hipStream_t>(stream));
    return check_launch("nz_superbee");
}

int nz_time_step(const nz_problem* p, const double* wave, double factor, double* dt, void* stream) {
    if (!p || !wave || !dt || !p->area) return fail(NZ_E_ARG, "null argument");
    const unsigned nb = static_cast<unsigned>((p->n_elem + nz::BS - 1) / nz::BS);
    hipLaunchKernelGGL(k_time_step, dim3(nb), dim3(nz::BS), 0, static_cast<hipStream_t>(stream),
                       wave, p->area, p->n_elem, factor, dt);
    return check_launch("nz_time_step");
}

}  // extern "C"
