// Mesh connectivity and per-element / per-edge geometry on the device
// (gfx950), behind nz_build_edges and nz_build_geometry in nozzle_hip.h.
//
// Connectivity follows the reference's build_edges
// (src/geometry/mesh.py:77-208) rule for rule, so the maps are bit-identical:
//   * every element face (slot = e * nface + f) is keyed by its corner pair,
//     key = min * n_nodes + max;
//   * a stable sort of (key, slot) puts the two faces of an interior edge
//     next to each other, lower slot first; interior edges are numbered in
//     key order and owned by the lower slot (so by the lower element);
//   * unpaired faces are boundary edges, tagged from the midpoint of their
//     corners in the tagger's coordinates (inflow x < tol, outflow
//     x > length - tol, axis |y| < tol, wall otherwise: mesh.py:340-358),
//     kept in key order within a tag and grouped by tag 1..4.
// The sorts are hipCUB radix sorts (stable), the numbering exclusive scans.
//
// Geometry follows build_operators (src/spatial/operators.py:209-342): the
// Jacobian and its inverse at the volume points, w * det, the element area
// and cell-average weights, the inverse mass matrix of non-affine elements
// (Cholesky of the SPD mass matrix), and per edge the owner-outward unit
// normal, |dX/dsigma| and the length.  It writes the element kernel's
// compact layout (nz_problem) and, when asked, the reference layout.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <cstdio>

#include "nozzle_hip.h"

namespace {

constexpr int BS = 256;
thread_local char g_merr[512] = "";

int mfail(int code, const char* msg) {
    std::snprintf(g_merr, sizeof g_merr, "%s", msg);
    return code;
}

int mcheck(hipError_t st, const char* what) {
    if (st == hipSuccess) st = hipGetLastError();
    if (st != hipSuccess) {
        std::snprintf(g_merr, sizeof g_merr, "%s: %s", what, hipGetErrorString(st));
        return NZ_E_LAUNCH;
    }
    return NZ_OK;
}

inline unsigned nblk(int64_t n) { return static_cast<unsigned>((n + BS - 1) / BS); }

inline int key_bits(int64_t n_nodes) {
    // keys run up to n_nodes^2 - 1
    const unsigned __int128 top = static_cast<unsigned __int128>(n_nodes) * n_nodes;
    int b = 1;
    while (b < 64 && (static_cast<unsigned __int128>(1) << b) < top) ++b;
    return b;
}

// ------------------------------------------------------------ connectivity
__global__ void k_face_keys(nz_mesh_desc m, uint64_t* __restrict__ key, uint32_t* __restrict__ slot,
                            int64_t n) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (s >= n) return;
    const int64_t e = s / m.nface;
    const int f = static_cast<int>(s - e * m.nface);
    const int64_t a = m.elem_nodes[e * m.nodes_per_elem + m.face_ends[2 * f]];
    const int64_t b = m.elem_nodes[e * m.nodes_per_elem + m.face_ends[2 * f + 1]];
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    key[s] = static_cast<uint64_t>(lo) * static_cast<uint64_t>(m.n_nodes) + static_cast<uint64_t>(hi);
    slot[s] = static_cast<uint32_t>(s);
}

// pair-start and boundary flags in key order; a key seen three times or more
// marks the mesh non-conforming
__global__ void k_pair_flags(const uint64_t* __restrict__ sk, int64_t n, int32_t* __restrict__ fi,
                             int32_t* __restrict__ fb, int64_t* __restrict__ counts) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (k >= n) return;
    const bool next = k + 1 < n && sk[k + 1] == sk[k];
    const bool prev = k > 0 && sk[k - 1] == sk[k];
    fi[k] = next ? 1 : 0;
    fb[k] = (next || prev) ? 0 : 1;
    if (next && prev) atomicAdd(reinterpret_cast<unsigned long long*>(&counts[6]), 1ull);
}

__global__ void k_counts(const int32_t* __restrict__ fi, const int32_t* __restrict__ fb,
                         const int32_t* __restrict__ ii, const int32_t* __restrict__ ib, int64_t n,
                         int64_t* __restrict__ counts) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) {
        counts[0] = ii[n - 1] + fi[n - 1];
        counts[1] = ib[n - 1] + fb[n - 1];
    }
}

__global__ void k_interior(nz_mesh_desc m, nz_edge_maps o, const uint32_t* __restrict__ sv,
                           const int32_t* __restrict__ fi, const int32_t* __restrict__ ii,
                           int64_t n) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (k >= n || !fi[k]) return;
    const int64_t i = ii[k];
    const int64_t s0 = sv[k], s1 = sv[k + 1];  // s0 < s1: the stable sort keeps slot order
    const int nf = m.nface;
    const int64_t e0 = s0 / nf, e1 = s1 / nf;
    const int f0 = static_cast<int>(s0 - e0 * nf), f1 = static_cast<int>(s1 - e1 * nf);
    o.iedge_elem[2 * i] = e0;
    o.iedge_elem[2 * i + 1] = e1;
    o.iedge_face[2 * i] = f0;
    o.iedge_face[2 * i + 1] = f1;
    o.face_edge[s0] = i;
    o.face_edge[s1] = i;
    o.face_side[s0] = 0;
    o.face_side[s1] = 1;
    o.face_neighbour[s0] = e1;
    o.face_neighbour[s1] = e0;
    if (o.edge_elem) {
        o.edge_elem[i] = e0;
        o.edge_face[i] = f0;
    }
    if (o.fe_nbr) {
        o.fe_nbr[s0] = static_cast<int32_t>(e1);
        o.fe_nbr[s1] = static_cast<int32_t>(e0);
        o.fe_info[s0] = static_cast<int32_t>((i << 3) | (0 << 2) | f1);
        o.fe_info[s1] = static_cast<int32_t>((i << 3) | (1 << 2) | f0);
    }
}

__device__ __forceinline__ int64_t tag_of(const nz_mesh_desc& m, int64_t s) {
    const int64_t e = s / m.nface;
    const int f = static_cast<int>(s - e * m.nface);
    const int64_t a = m.elem_nodes[e * m.nodes_per_elem + m.face_ends[2 * f]];
    const int64_t b = m.elem_nodes[e * m.nodes_per_elem + m.face_ends[2 * f + 1]];
    const double x = 0.5 * (m.tag_coords[2 * a] + m.tag_coords[2 * b]);
    const double y = 0.5 * (m.tag_coords[2 * a + 1] + m.tag_coords[2 * b + 1]);
    // later tests win, as the reference tagger assigns them in this order
    int64_t t = NZ_TAG_WALL;
    if (fabs(y) < m.tag_tol) t = NZ_TAG_AXIS;
    if (x < m.tag_tol) t = NZ_TAG_INFLOW;
    if (x > m.tag_length - m.tag_tol) t = NZ_TAG_OUTFLOW;
    return t;
}

__global__ void k_boundary_tags(nz_mesh_desc m, const uint32_t* __restrict__ sv,
                                const int32_t* __restrict__ fb, const int32_t* __restrict__ ib,
                                int64_t n, uint32_t* __restrict__ btag, uint32_t* __restrict__ bslot,
                                int64_t* __restrict__ counts) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (k >= n || !fb[k]) return;
    const int64_t j = ib[k];
    const uint32_t s = sv[k];
    const int64_t t = tag_of(m, s);
    btag[j] = static_cast<uint32_t>(t);
    bslot[j] = s;
    atomicAdd(reinterpret_cast<unsigned long long*>(&counts[1 + t]), 1ull);
}

__global__ void k_boundary_out(nz_mesh_desc m, nz_edge_maps o, const uint32_t* __restrict__ btag,
                               const uint32_t* __restrict__ bslot, const int64_t* __restrict__ counts) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    const int64_t nb = counts[1], ni = counts[0];
    if (j >= nb) return;
    const int64_t s = bslot[j];
    const int nf = m.nface;
    const int64_t e = s / nf;
    const int f = static_cast<int>(s - e * nf);
    const int64_t t = btag[j];
    o.bedge_elem[j] = e;
    o.bedge_face[j] = f;
    o.bedge_tag[j] = t;
    o.face_edge[s] = ni + j;
    o.face_side[s] = 0;
    o.face_neighbour[s] = e;
    if (o.edge_elem) {
        o.edge_elem[ni + j] = e;
        o.edge_face[ni + j] = f;
    }
    if (o.fe_nbr) {
        o.fe_nbr[s] = -static_cast<int32_t>(t);
        o.fe_info[s] = static_cast<int32_t>(((ni + j) << 3) | f);
        o.bedge[2 * j] = static_cast<int32_t>(e);
        o.bedge[2 * j + 1] = static_cast<int32_t>((f << 3) | t);
    }
}

// boundary-tag keys past the boundary count: above every tag, so they sort last
__global__ void k_tail_tags(uint32_t* __restrict__ bt, const int64_t* __restrict__ counts,
                            int64_t n) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (j < n && j >= counts[1]) bt[j] = 7u;
}

// workspace carve-up: every region 256-byte aligned
struct EdgeWork {
    uint64_t *key_in, *key_out;
    uint32_t *slot_in, *slot_out;
    int32_t *fi, *fb, *ii, *ib;
    uint32_t *btag_in, *btag_out, *bslot_in, *bslot_out;
    void* cub;
    size_t cub_bytes;
    size_t total;
};

inline size_t up(size_t b) { return (b + 255) & ~size_t(255); }

size_t cub_bytes_for(int64_t n) {
    size_t a = 0, b = 0, c = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, static_cast<const uint64_t*>(nullptr),
                                       static_cast<uint64_t*>(nullptr),
                                       static_cast<const uint32_t*>(nullptr),
                                       static_cast<uint32_t*>(nullptr), n);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<const uint32_t*>(nullptr),
                                       static_cast<uint32_t*>(nullptr),
                                       static_cast<const uint32_t*>(nullptr),
                                       static_cast<uint32_t*>(nullptr), n);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, static_cast<const int32_t*>(nullptr),
                                     static_cast<int32_t*>(nullptr), n);
    size_t m = a > b ? a : b;
    return m > c ? m : c;
}

EdgeWork carve(void* base, int64_t n) {
    EdgeWork w{};
    char* p = static_cast<char*>(base);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* q = p ? p + off : nullptr;
        off += up(bytes);
        return q;
    };
    w.key_in = reinterpret_cast<uint64_t*>(take(8 * n));
    w.key_out = reinterpret_cast<uint64_t*>(take(8 * n));
    w.slot_in = reinterpret_cast<uint32_t*>(take(4 * n));
    w.slot_out = reinterpret_cast<uint32_t*>(take(4 * n));
    w.fi = reinterpret_cast<int32_t*>(take(4 * n));
    w.fb = reinterpret_cast<int32_t*>(take(4 * n));
    w.ii = reinterpret_cast<int32_t*>(take(4 * n));
    w.ib = reinterpret_cast<int32_t*>(take(4 * n));
    w.btag_in = reinterpret_cast<uint32_t*>(take(4 * n));
    w.btag_out = reinterpret_cast<uint32_t*>(take(4 * n));
    w.bslot_in = reinterpret_cast<uint32_t*>(take(4 * n));
    w.bslot_out = reinterpret_cast<uint32_t*>(take(4 * n));
    w.cub_bytes = cub_bytes_for(n);
    w.cub = take(w.cub_bytes);
    w.total = off;
    return w;
}

// -------------------------------------------------------------- geometry
// One thread per (element, volume point): Jacobian, its inverse, w * det,
// physical point (src/spatial/operators.py:245-268, element_metrics).
__global__ void k_geo_points(nz_geometry_desc g, nz_geometry_out o) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (it >= g.n_elem * g.nqv) return;
    const int64_t e = it / g.nqv;
    const int q = static_cast<int>(it - e * g.nqv);
    const double* T = g.tables;
    double xxi = 0.0, xeta = 0.0, yxi = 0.0, yeta = 0.0, px = 0.0, py = 0.0;
    for (int n = 0; n < g.nq; ++n) {
        const int64_t node = g.elem_nodes[e * g.nq + n];
        const double x = g.coords[2 * node], y = g.coords[2 * node + 1];
        const double dxi = T[g.off.dgeom_dxi_vol + n * g.nqv + q];
        const double deta = T[g.off.dgeom_deta_vol + n * g.nqv + q];
        const double gv = T[g.off.geom_vol + n * g.nqv + q];
        xxi += x * dxi;
        xeta += x * deta;
        yxi += y * dxi;
        yeta += y * deta;
        px += gv * x;
        py += gv * y;
    }
    const double det = xxi * yeta - xeta * yxi;
    const double inv = 1.0 / det;
    const double m[4] = {yeta * inv, -xeta * inv, -yxi * inv, xxi * inv};
    const double wdet = T[g.off.w_vol + q] * det;
    if (!(det > 0.0)) atomicAdd(reinterpret_cast<unsigned long long*>(o.bad_jacobian), 1ull);
    if (g.affine) {
        if (q == 0) {
            double* ge = o.geo + e * 5;
            for (int c = 0; c < 4; ++c) ge[c] = m[c];
            ge[4] = det;
        }
    } else {
        double* ge = o.geo + (e * g.nqv + q) * 5;
        for (int c = 0; c < 4; ++c) ge[c] = m[c];
        ge[4] = wdet;
    }
    if (o.det) o.det[e * g.nqv + q] = det;
    if (o.metrics)
        for (int c = 0; c < 4; ++c) o.metrics[(e * g.nqv + q) * 4 + c] = m[c];
    if (o.wdet) o.wdet[e * g.nqv + q] = wdet;
    if (o.xy_vol) {
        o.xy_vol[(e * g.nqv + q) * 2] = px;
        o.xy_vol[(e * g.nqv + q) * 2 + 1] = py;
    }
}

// One thread per element: area and cell-average weights from w * det at the
// points (operators.py: area = sum_q w det, mean_w = phi (w det) / area).
__global__ void k_geo_elements(nz_geometry_desc g, nz_geometry_out o) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (e >= g.n_elem) return;
    const double* T = g.tables;
    auto wdet = [&](int q) -> double {
        if (o.wdet) return o.wdet[e * g.nqv + q];
        if (g.affine) return T[g.off.w_vol + q] * o.geo[e * 5 + 4];
        return o.geo[(e * g.nqv + q) * 5 + 4];
    };
    double area = 0.0;
    for (int q = 0; q < g.nqv; ++q) area += wdet(q);
    o.area[e] = area;
    double* mw = o.mean_w_full ? o.mean_w_full + e * g.nbf : (g.affine ? nullptr : o.mean_w + e * g.nbf);
    if (mw) {
        for (int i = 0; i < g.nbf; ++i) {
            double s = 0.0;
            for (int q = 0; q < g.nqv; ++q) s += T[g.off.phi_vol + i * g.nqv + q] * wdet(q);
            mw[i] = s / area;
        }
        if (o.mean_w_full && !g.affine)
            for (int i = 0; i < g.nbf; ++i) o.mean_w[e * g.nbf + i] = mw[i];
    }
}

// One thread per element (non-affine): the mass matrix M_ij = sum_q phi_i
// phi_j w det, its Cholesky factor and the inverse, stored as the upper
// triangle row-major (the element kernel's packed layout).
template <int N>
__global__ __launch_bounds__(64) void k_geo_minv(nz_geometry_desc g, nz_geometry_out o) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (e >= g.n_elem) return;
    const double* T = g.tables;
    double L[N * N];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int q = 0; q < g.nqv; ++q)
                s += T[g.off.phi_vol + i * g.nqv + q] * T[g.off.phi_vol + j * g.nqv + q] *
                     o.geo[(e * g.nqv + q) * 5 + 4];
            L[i * N + j] = s;
        }
    // in-place Cholesky: lower triangle of L
    for (int j = 0; j < N; ++j) {
        double d = L[j * N + j];
        for (int k = 0; k < j; ++k) d -= L[j * N + k] * L[j * N + k];
        d = sqrt(d);
        L[j * N + j] = d;
        for (int i = j + 1; i < N; ++i) {
            double s = L[i * N + j];
            for (int k = 0; k < j; ++k) s -= L[i * N + k] * L[j * N + k];
            L[i * N + j] = s / d;
        }
    }
    // W = L^-1 (lower), in the upper triangle's mirror: Wt[j][i] = W[i][j]
    double W[N * N];
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s -= L[i * N + k] * W[k * N + j];
            W[i * N + j] = s / L[i * N + i];
        }
        W[i * N + i] = 1.0 / L[i * N + i];
        for (int j = i + 1; j < N; ++j) W[i * N + j] = 0.0;
    }
    // M^-1 = W^T W; entry (i, j), i <= j: sum_{k >= j} W[k][i] W[k][j]
    double* out = o.minv + e * (N * (N + 1) / 2);
    int pos = 0;
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) {
            double s = 0.0;
            for (int k = j; k < N; ++k) s += W[k * N + i] * W[k * N + j];
            out[pos++] = s;
        }
}

// One thread per edge: owner-outward unit normal, |dX/dsigma| and physical
// point at each face point, and the length (operators.py edge_metrics).
__global__ void k_geo_edges(nz_geometry_desc g, nz_geometry_out o) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * BS + threadIdx.x;
    if (k >= g.n_edge) return;
    const int64_t e = g.edge_elem[k];
    const int f = static_cast<int>(g.edge_face[k]);
    const double* T = g.tables;
    double len = 0.0;
    for (int q = 0; q < g.nqf; ++q) {
        double tx = 0.0, ty = 0.0, px = 0.0, py = 0.0;
        for (int n = 0; n < g.nq; ++n) {
            const int64_t node = g.elem_nodes[e * g.nq + n];
            const double x = g.coords[2 * node], y = g.coords[2 * node + 1];
            const double d = T[g.off.dgeom_dsigma_face + (f * g.nq + n) * g.nqf + q];
            const double gv = T[g.off.geom_face + (f * g.nq + n) * g.nqf + q];
            tx += d * x;
            ty += d * y;
            px += gv * x;
            py += gv * y;
        }
        const double jac = sqrt(tx * tx + ty * ty);
        const double nx = ty / jac, ny = -tx / jac;
        len += jac * T[g.off.w_face + q];
        if (!g.straight || q == 0) {
            double* eg = g.straight ? o.edge_geo + k * 3 : o.edge_geo + (k * g.nqf + q) * 3;
            eg[0] = nx;
            eg[1] = ny;
            eg[2] = jac;
        }
        if (o.normal) {
            o.normal[(k * g.nqf + q) * 2] = nx;
            o.normal[(k * g.nqf + q) * 2 + 1] = ny;
        }
        if (o.jac) o.jac[k * g.nqf + q] = jac;
        if (o.xy_face) {
            o.xy_face[(k * g.nqf + q) * 2] = px;
            o.xy_face[(k * g.nqf + q) * 2 + 1] = py;
        }
    }
    if (o.length) o.length[k] = len;
}

}  // namespace

extern "C" {

const char* nz_mesh_last_error(void) { return g_merr; }

int64_t nz_build_edges_workspace_bytes(int64_t n_elem, int32_t nface) {
    if (n_elem <= 0 || nface < 3 || nface > 4) return -1;
    const int64_t n = n_elem * nface;
    return static_cast<int64_t>(carve(nullptr, n).total);
}

int nz_build_edges(const nz_mesh_desc* m, const nz_edge_maps* o, void* work, int64_t work_bytes,
                   void* stream) {
    if (!m || !o) return mfail(NZ_E_ARG, "null argument");
    if (m->n_elem <= 0 || m->nface < 3 || m->nface > 4 || m->nodes_per_elem <= 0 ||
        m->n_nodes <= 0)
        return mfail(NZ_E_ARG, "invalid mesh description");
    const int64_t n = m->n_elem * m->nface;
    if (n >= (int64_t{1} << 31) || m->n_nodes > (int64_t{1} << 31))
        return mfail(NZ_E_ARG, "mesh too large for 32-bit slots");
    for (int f = 0; f < 2 * m->nface; ++f)
        if (m->face_ends[f] < 0 || m->face_ends[f] >= m->nodes_per_elem)
            return mfail(NZ_E_ARG, "face_ends out of range");
    if (!m->elem_nodes || !m->tag_coords || !o->iedge_elem || !o->iedge_face || !o->bedge_elem ||
        !o->bedge_face || !o->bedge_tag || !o->face_edge || !o->face_side || !o->face_neighbour ||
        !o->counts)
        return mfail(NZ_E_ARG, "null array");
    if ((o->fe_nbr || o->fe_info || o->bedge) && !(o->fe_nbr && o->fe_info && o->bedge))
        return mfail(NZ_E_ARG, "fe_nbr, fe_info and bedge go together");
    if ((o->edge_elem != nullptr) != (o->edge_face != nullptr))
        return mfail(NZ_E_ARG, "edge_elem and edge_face go together");
    EdgeWork w = carve(work, n);
    if (!work || work_bytes < static_cast<int64_t>(w.total))
        return mfail(NZ_E_ARG, "workspace smaller than nz_build_edges_workspace_bytes");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = mcheck(hipMemsetAsync(o->counts, 0, 8 * sizeof(int64_t), st), "counts");
    if (rc) return rc;
    hipLaunchKernelGGL(k_face_keys, dim3(nblk(n)), dim3(BS), 0, st, *m, w.key_in, w.slot_in, n);
    size_t cb = w.cub_bytes;
    rc = mcheck(hipcub::DeviceRadixSort::SortPairs(w.cub, cb, w.key_in, w.key_out, w.slot_in,
                                                    w.slot_out, n, 0, key_bits(m->n_nodes), st),
                "face key sort");
    if (rc) return rc;
    hipLaunchKernelGGL(k_pair_flags, dim3(nblk(n)), dim3(BS), 0, st, w.key_out, n, w.fi, w.fb,
                       o->counts);
    cb = w.cub_bytes;
    rc = mcheck(hipcub::DeviceScan::ExclusiveSum(w.cub, cb, w.fi, w.ii, n, st), "interior scan");
    if (rc) return rc;
    cb = w.cub_bytes;
    rc = mcheck(hipcub::DeviceScan::ExclusiveSum(w.cub, cb, w.fb, w.ib, n, st), "boundary scan");
    if (rc) return rc;
    hipLaunchKernelGGL(k_counts, dim3(1), dim3(64), 0, st, w.fi, w.fb, w.ii, w.ib, n, o->counts);
    hipLaunchKernelGGL(k_interior, dim3(nblk(n)), dim3(BS), 0, st, *m, *o, w.slot_out, w.fi, w.ii,
                       n);
    hipLaunchKernelGGL(k_boundary_tags, dim3(nblk(n)), dim3(BS), 0, st, *m, w.slot_out, w.fb, w.ib,
                       n, w.btag_in, w.bslot_in, o->counts);
    // The boundary count is on the device: sort the whole slot range; entries
    // past the count hold keys above every tag (filled here) and sort last.
    hipLaunchKernelGGL(k_tail_tags, dim3(nblk(n)), dim3(BS), 0, st, w.btag_in, o->counts, n);
    cb = w.cub_bytes;
    rc = mcheck(hipcub::DeviceRadixSort::SortPairs(w.cub, cb, w.btag_in, w.btag_out, w.bslot_in,
                                                    w.bslot_out, n, 0, 3, st),
                "boundary tag sort");
    if (rc) return rc;
    hipLaunchKernelGGL(k_boundary_out, dim3(nblk(n)), dim3(BS), 0, st, *m, *o, w.btag_out,
                       w.bslot_out, o->counts);
    return mcheck(hipSuccess, "nz_build_edges");
}

int nz_build_geometry(const nz_geometry_desc* g, const nz_geometry_out* o, void* stream) {
    if (!g || !o) return mfail(NZ_E_ARG, "null argument");
    if (g->n_elem <= 0 || g->nqv <= 0 || g->nq <= 0 || g->nbf <= 0 || g->nqf <= 0 ||
        g->n_edge < 0)
        return mfail(NZ_E_ARG, "invalid geometry description");
    if (!g->elem_nodes || !g->coords || !g->tables || !o->geo || !o->area || !o->edge_geo ||
        !o->bad_jacobian || (g->n_edge > 0 && (!g->edge_elem || !g->edge_face)))
        return mfail(NZ_E_ARG, "null array");
    if (!g->affine && (!o->minv || !o->mean_w))
        return mfail(NZ_E_ARG, "non-affine elements need minv and mean_w");
    const int N = g->nbf;
    if (!g->affine && N != 1 && N != 3 && N != 4 && N != 6 && N != 9 && N != 10 && N != 16)
        return mfail(NZ_E_ARG, "unsupported basis size for the inverse mass");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = mcheck(hipMemsetAsync(o->bad_jacobian, 0, sizeof(int64_t), st), "bad_jacobian");
    if (rc) return rc;
    hipLaunchKernelGGL(k_geo_points, dim3(nblk(g->n_elem * g->nqv)), dim3(BS), 0, st, *g, *o);
    hipLaunchKernelGGL(k_geo_elements, dim3(nblk(g->n_elem)), dim3(BS), 0, st, *g, *o);
    if (!g->affine) {
        const dim3 gr(static_cast<unsigned>((g->n_elem + 63) / 64));
        switch (N) {
            case 1: hipLaunchKernelGGL(k_geo_minv<1>, gr, dim3(64), 0, st, *g, *o); break;
            case 3: hipLaunchKernelGGL(k_geo_minv<3>, gr, dim3(64), 0, st, *g, *o); break;
            case 4: hipLaunchKernelGGL(k_geo_minv<4>, gr, dim3(64), 0, st, *g, *o); break;
            case 6: hipLaunchKernelGGL(k_geo_minv<6>, gr, dim3(64), 0, st, *g, *o); break;
            case 9: hipLaunchKernelGGL(k_geo_minv<9>, gr, dim3(64), 0, st, *g, *o); break;
            case 10: hipLaunchKernelGGL(k_geo_minv<10>, gr, dim3(64), 0, st, *g, *o); break;
            default: hipLaunchKernelGGL(k_geo_minv<16>, gr, dim3(64), 0, st, *g, *o); break;
        }
    }
    if (g->n_edge > 0)
        hipLaunchKernelGGL(k_geo_edges, dim3(nblk(g->n_edge)), dim3(BS), 0, st, *g, *o);
    return mcheck(hipSuccess, "nz_build_geometry");
}

}  // extern "C"
