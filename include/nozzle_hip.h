/*
 * C ABI of the MI355X DG residual / SSP-RK3 engine (gfx950).
 *
 * Plain pointers and sizes only.  Every pointer argument below is a DEVICE
 * pointer unless stated otherwise; `stream` is a hipStream_t passed as void*.
 * Calls are asynchronous on `stream`; return codes are NZ_OK or a negative
 * NZ_E* value, with a message from nz_last_error() (thread-local).  No C++
 * exception crosses this boundary.  The library allocates nothing: every
 * buffer, including scratch, is passed in by the caller.
 *
 * Arrays are FP64 and C-contiguous.  The state U is (E, N, 4) in the
 * reference's element-batched layout (src/spatial/operators.py:138-190),
 * N = n_basis, with conserved variables (rho, rho*vx, rho*vy, rho*E).
 *
 * The Python backend (2d_nozzle_amd/backend.py, class HipBackend) drives
 * these entry points behind the reference's Backend interface
 * (src/backends/base.py:81-201).
 */
#ifndef NOZZLE_HIP_H
#define NOZZLE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NZ_OK 0
#define NZ_E_ARG -1       /* invalid argument or unsupported (kind, p, Q) */
#define NZ_E_LAUNCH -2    /* kernel launch failed */

/* Offsets (in doubles) of the reference tables inside nz_problem.tables. */
typedef struct {
    int32_t phi_vol;   /* (N, nqv)                basis at volume points          */
    int32_t dxi;       /* (N, nqv)                d phi / d xi                    */
    int32_t deta;      /* (N, nqv)                d phi / d eta                   */
    int32_t axi;       /* (N, nqv)  sum_j Minv_ref[i,j] w_q dphi_j/dxi (affine rate) */
    int32_t aeta;      /* (N, nqv)                                                */
    int32_t phi_face;  /* (2, nface, N, nqf)      side 0 walks sigma, 1 walks 1-sigma */
    int32_t lift;      /* (2, nface, N, nqf)  Minv_ref @ phi_face (affine rate)   */
    int32_t w_vol;     /* (nqv,)                                                  */
    int32_t w_face;    /* (nqf,)                                                  */
    int32_t mean_w;    /* (N,)  cell-average weights of an affine element         */
    int32_t minv_ref;  /* (N, N) inverse reference mass matrix (affine)           */
    int32_t vol_t;     /* (nqv, 5, N) point-major copies: phi, dxi, deta, axi, aeta */
} nz_table_offsets;

/* Everything one mesh/order/operating point needs.  Filled on the host by
 * HipBackend; pointers are device pointers owned by the caller. */
typedef struct {
    int64_t n_elem;
    int64_t n_edge;
    int32_t nface, nbf, nqv, nqf;
    int32_t affine;     /* 1: straight-sided triangles, constant metrics per element */
    int32_t straight;   /* 1: straight edges (Q = 1), one normal/jacobian per edge   */
    /* geometry */
    const double *geo;      /* affine: (E, 5) = xi_x, xi_y, eta_x, eta_y, det
                               else:   (E, nqv, 5) = xi_x, xi_y, eta_x, eta_y, w*det */
    const double *area;     /* (E,) element area, sum_q w_q det_q                   */
    const double *minv;     /* (E, N(N+1)/2) inverse mass matrices (non-affine only):
                             * upper triangles, row-major, entry (i, j) for i <= j */
    const double *mean_w;   /* (E, N) cell-average weights (non-affine only)        */
    const double *edge_geo; /* straight: (F, 3) = nx, ny, |dX/dsigma|;  else (F, nqf, 3) */
    const int32_t *fe_nbr;  /* (E, nface) element across the face, or -tag (1..4)   */
    const int32_t *fe_info; /* (E, nface) (edge << 3) | (side << 2) | other_face    */
    int64_t n_int;          /* interior edges; boundary edge b has global id n_int + b */
    const int32_t *bedge;   /* (Fb, 2): owner element, (local face << 3) | tag      */
    double *bflux;          /* (Fb, nqf, 5) workspace: weighted boundary flux, speed */
    const double *tables;   /* reference tables, laid out by `off`                  */
    nz_table_offsets off;
    /* flow (src/config.py:43-180; src/backends/numba_backend.py:69-109) */
    double gamma, efix, low_mach;
    int32_t flux_id;        /* 0 roe, 1 hllc, 2 ausmup */
    int32_t limiter;        /* 0 none, 1 positivity (fused), 2 superbee then positivity (separate passes) */
    double rho_t_bf, p_t_bf, band_bf;
    double at2, at, rho_t, ca, sa, p_back;
    /* limiter (src/spatial/limiter.py:98-243) */
    double rho_floor, p_floor, fraction, lebesgue;
    int32_t bisect_steps;
    /* boundary fluxes carried from one Runge-Kutta stage to the next
     * (nz_stage only; nz_residual, nz_rate and nz_rate_noalloc refuse
     * a nonzero value).  Bit 0: `bflux` already holds the
     * boundary fluxes of Uin, written by the previous nz_stage on this
     * problem, so the boundary pre-pass is skipped.  Bit 1: after Uout is
     * complete, each workgroup writes the boundary fluxes of Uout for its own
     * boundary faces into `bflux`, for the next stage (the values the
     * pre-pass would compute, bit for bit).  Only valid when nothing changes
     * Uout between the two stages: not with limiter 2, whose passes run
     * after the stage.  No counterpart in the reference, which evaluates the
     * boundary faces in every residual (src/backends/_numba_kernels.py). */
    int32_t bflux_chain;
    double wnorm[4];        /* residual-norm component weights (src/backends/base.py:42-78) */
    /* optional tile table for the element kernel of nz_residual / nz_rate /
     * nz_stage, one workgroup per entry in this order: entry k covers
     * elements [tiles[3k], tiles[3k+1]), at most 16 of them, and writes its
     * norm partial to scratch[tiles[3k+2]] (slots distinct, < n_tiles).
     * NULL: consecutive tiles of 16 over the whole mesh, slot = tile index.
     * Used to order the launch for cache locality, to march several disjoint
     * meshes (designs) in one launch and to leave converged ones out. */
    const int64_t *tiles;
    int64_t n_tiles;
    /* optional tile table for the adjoint kernels of nz_residual_vjp,
     * nz_adjoint_rk4_stage and nz_adjoint_defect, one workgroup per entry:
     * entry k covers elements [vjp_tiles[2k], vjp_tiles[2k+1]), at most
     * nz_adjoint_tile_elements(p) of them.  Elements outside every entry are
     * not touched (their outputs keep what they held); nz_adjoint_defect
     * sums over the listed elements only.  The boundary pre-pass still
     * covers every boundary edge.  NULL: consecutive tiles over the whole
     * mesh.  Used to leave converged designs of a batch out of the adjoint
     * march. */
    const int64_t *vjp_tiles;
    int64_t n_vjp_tiles;
    /* the state-cotangent VJP kernel: 0 one-wave tiles of 16 elements (the
     * element kernel's mapping), 1 four-wave tiles of 64 (or 128 when
     * N <= 4); the two agree to round-off.  HipBackend picks it per (element
     * kind, order, geometry order, flux) by timing both. */
    int32_t vjp_kernel;
    /* 1 when `tiles` lists every consecutive tile of nz::TE (16) elements
     * exactly once, with norm slot = the tile's index (HipBackend's locality
     * order): the one-wave adjoint launches then take their tiles from it as
     * well (neighbour gathers of a spatially coherent launch hit L2).  0 for
     * any other table (a batch's live designs). */
    int32_t tiles_cover_mesh;
    /* optional patch table for the element kernel of nz_residual / nz_rate /
     * nz_stage on affine (straight-sided triangle) meshes, used in place of
     * `tiles` when set: one workgroup per patch, n_patches of them.  Patch k
     * is the 18 int32 at patches[18k]: up to 16 element ids (the first
     * `count` valid, the rest -1), count (1..16), norm slot (distinct,
     * < n_patches).  Every element is in exactly one patch.  Patches of
     * elements that share edges solve each shared edge's flux once; results
     * are bit-identical to the tiled kernel's.  No counterpart in the
     * reference (its edge pass solves every interior edge once by
     * construction, src/backends/_numba_kernels.py:484-583). */
    const int32_t *patches;
    int64_t n_patches;
    /* optional, for a march of state cotangents at a fixed U (the adjoint):
     * the boundary-flux Jacobians of that U from nz_boundary_jacobian.  When
     * set, nz_residual_vjp, nz_adjoint_rk4_stage and nz_adjoint_defect on
     * one-wave tiles (vjp_kernel 0) without geometry cotangents skip their
     * boundary pre-pass and apply the Jacobian to the cotangent's traces
     * instead; every U passed must then be the one it was formed from.  No
     * counterpart in the reference, whose jax.vjp re-linearises the boundary
     * fluxes at every call (src/sensitivity.py:289-383). */
    const double *bjac;
} nz_problem;

/* Number of scratch doubles nz_* calls need for norm partials. */
int64_t nz_scratch_doubles(const nz_problem *p);

/* Elements per workgroup of the state-cotangent adjoint launches: the
 * widest entry a vjp_tiles table may hold (16 with vjp_kernel 0, else
 * the four-wave tile of 64, or 128 when N <= 4).  No
 * counterpart in the reference (its adjoint is a JAX vjp over the whole mesh,
 * src/sensitivity.py:317). */
int64_t nz_adjoint_tile_elements(const nz_problem *p);

/* Lagrange node j (0 <= j <= p) on local face f of the reference element
 * (nface 3: triangles, 4: quadrilaterals) of order p, in the reference's
 * node numbering (src/geometry/elements.py:57-75); -1 for arguments out of
 * range.  The basis functions of all other nodes vanish on that face, so the
 * element kernels form face traces from these p + 1 coefficients alone
 * (triangles, and quadrilaterals up to p = 2).  A host-side query with no
 * counterpart in the reference: tests pin it against the reference's face
 * tables. */
int nz_face_node(int nface, int p, int f, int j);

/* R(U) and the wave-speed sum, unfolded.
 * Replaces edge_pass + element_pass(apply_mass=False)
 * (src/backends/_numba_kernels.py:484-583, 667-759) and the NumPy
 * assembly.residual (src/spatial/assembly.py:153-157).  R: (E,N,4); wave: (E,). */
int nz_residual(const nz_problem *p, const double *U, double *R, double *wave, void *stream);

/* The rate -M^-1 R(U) and the wave-speed sum.
 * Replaces edge_pass + element_rate (src/backends/_numba_kernels.py:586-664). */
int nz_rate(const nz_problem *p, const double *U, double *F, double *wave, void *stream);

/* nz_rate without a caller-provided workspace: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The boundary pre-pass output, for which nz_rate uses
 * p->bflux, goes to a static device array of 2^20 doubles (8 MiB) per GPU,
 * held until the process ends; p->bflux is ignored.  Meshes with more than
 * nz_rate_noalloc_max_boundary_points() boundary face points (boundary edges
 * x nqf) are rejected with NZ_E_ARG before any launch.  Nothing here calls
 * hipMalloc or hipFree.  All calls on one GPU share that array: calls that
 * overlap on one GPU (different streams or threads) race; calls in stream
 * order do not.  Results are those of nz_rate, bit for bit. */
int nz_rate_noalloc(const nz_problem *p, const double *U, double *F, double *wave, void *stream);
int64_t nz_rate_noalloc_max_boundary_points(void);

/* out = M^-1 R per element (src/backends/_numba_kernels.py:762-777). */
int nz_inverse_mass(const nz_problem *p, const double *R, double *out, void *stream);

/* L2 projection onto the basis: U = M^-1 sum_q phi(q) (w det)_q values_q per
 * element.  Replaces project (src/timestepping/initialize.py:41-46), the step
 * that forms the initial state.  values: (E, nqv, 4) at the volume points;
 * U: (E, N, 4). */
int nz_project(const nz_problem *p, const double *values, double *U, void *stream);

/* Boundary integrals of the owner-side trace behind the observables
 * (thrust, mass flow, exit Mach).  Replaces the sums of performance,
 * axial_momentum_flux and mass_flux (src/post/postprocess.py:56-75,
 * 155-171, 193-226).  out (8 doubles): axial momentum flux through the
 * inflow and the outflow, -p n_x ds over the walls, mass flux through the
 * inflow and the outflow, the outflow length, and the integrals of the Mach
 * number and of p over the outflow (one half-channel; the observables take
 * twice the fluxes). */
int nz_boundary_integrals(const nz_problem *p, const double *U, double *out, void *stream);

/* One fused Runge-Kutta stage: rate of Uin, stage combination, positivity
 * limiter, written to Uout (which must not alias Uin or Un).
 *   first != 0:  Uout = L(Uin + dt * F(Uin)); dt is computed from the wave
 *                speeds with `dt_factor` and written to `dt`; the weighted
 *                residual-norm partial sums of F are written to `scratch`.
 *   first == 0:  Uout = L(a * Un + b * (Uin + dt * F(Uin))), dt read.
 * flags (nullable): per-element limiter activity bits (bit 0 mean repaired,
 * bit 1 probe bound violated).  counters (nullable, 2 x uint64, accumulated):
 * elements scaled, cell means repaired.
 * Replaces one stage of NumbaBackend._ssprk3_fused / _euler_fused
 * (src/backends/numba_backend.py:305-337): _rate_into, _dt_into,
 * rms_weighted, stage / combine_ssp and positivity_limit. */
int nz_stage(const nz_problem *p, const double *Uin, const double *Un, double *Uout,
             double *dt, double dt_factor, double a, double b, int first,
             double *scratch, uint8_t *flags, uint64_t *counters, void *stream);

/* Finish the residual norm from the partials nz_stage(first=1) left in
 * scratch: out[0] = sqrt(sum / (E*N*4)) (src/backends/_numba_kernels.py:789-814). */
int nz_norm_finish(const nz_problem *p, const double *scratch, double *out, void *stream);

/* Per-segment residual norms from the partials nz_stage(first=1) left in
 * `scratch` under a tile table: out[d] = sqrt(sum of scratch[seg[d] ..
 * seg[d+1]) / counts[d]), summed in the same order as nz_norm_finish sums a
 * whole mesh, so a design marched in a batch sees the norm it would see
 * alone.  (Batched form of src/backends/_numba_kernels.py:789-814.) */
int nz_norm_segments(const nz_problem *p, const double *scratch, const int64_t *seg,
                     const double *counts, int64_t n_seg, double *out, void *stream);

/* Weighted RMS norm of an arbitrary (E,N,4) array (Backend.norm,
 * src/backends/numpy_backend.py:55-60).  Uses scratch; result in out[0]. */
int nz_norm(const nz_problem *p, const double *A, double *scratch, double *out, void *stream);

/* In-place Zhang-Shu positivity limiter (src/backends/_numba_kernels.py:927-1093,
 * src/spatial/limiter.py:181-243).  flags/counters as in nz_stage. */
int nz_limit(const nz_problem *p, double *U, uint8_t *flags, uint64_t *counters, void *stream);

/* Superbee TVB slope limiter, in place, without the positivity pass that
 * apply_limiter runs after it (src/spatial/limiter.py:286-405,408-429).
 * area_total: sum of element areas; work: nz_superbee_work_doubles(p)
 * doubles; counters[0] (nullable) += elements whose slope was scaled.
 * Elements with theta = 1 are left bit-identical. */
int64_t nz_superbee_work_doubles(const nz_problem *p);
int nz_superbee(const nz_problem *p, double *U, double tvb, double area_total, double *work,
                uint64_t *counters, void *stream);

/* ---- adjoint (src/sensitivity.py:289-460) ---------------------------------
 * Workspace of the boundary-face pre-pass: nz_adjoint_work_doubles(p). */
int64_t nz_adjoint_work_doubles(const nz_problem *p);

/* dU = (dR/dU)^T lam for the unfolded residual R at state U (the JAX vjp of
 * assembly.residual, src/sensitivity.py:313-317), with flux, boundary and
 * Euler-flux Jacobians recomputed per point.  Optional cotangents (nullable):
 *   gmet  (E, nqv, 5): d(lam.R)/d(xi_x, xi_y, eta_x, eta_y, w*det) per volume point;
 *   gedge (F, nqf, 3): d(lam.R)/d(n_x, n_y, |dX/dsigma|) per edge point;
 *   gpb   (Fb, nqf):   d(lam.R)/d(back pressure) per boundary point.
 * They chain to node coordinates and design variables on the host
 * (src/sensitivity.py:453-457). */
int nz_residual_vjp(const nz_problem *p, const double *U, const double *lam, double *dU,
                    double *gmet, double *gedge, double *gpb, double *work, void *stream);

/* One stage (1..4) of the classical RK4 pseudo-time march of the adjoint,
 * rate(lam) = -M^-1 (vjp(lam) - seed) with the frozen step dt
 * (src/sensitivity.py:330-357):
 *   stage 1: acc = k; lam_next = base + dt/2 k     stage 3: acc += 2k; lam_next = base + dt k
 *   stage 2: acc += 2k; lam_next = base + dt/2 k   stage 4: lam_next = base + dt/6 (acc + k) */
int nz_adjoint_rk4_stage(const nz_problem *p, const double *U, const double *lam_in,
                         const double *lam_base, const double *seed, const double *dt,
                         double *acc, double *lam_next, int stage, double *work, void *stream);

/* out[0] = sqrt(mean((vjp(lam) - seed)^2)) (src/sensitivity.py:378). */
int nz_adjoint_defect(const nz_problem *p, const double *U, const double *lam, const double *seed,
                      double *work, double *scratch, double *out, void *stream);

/* The boundary-flux Jacobians of U for nz_problem.bjac: per boundary face
 * point (Fb, nqf), 20 doubles: J[s][d] = dFb_s / du_d (row-major, 16) at the
 * owner's trace, the point's |dX/dsigma| w_q, and three zeros.  The Jacobian
 * is the one the boundary pre-pass of nz_residual_vjp differentiates
 * (src/spatial/boundary.py:41-214 through jax.vjp, src/sensitivity.py);
 * no reference function computes it on its own. */
int nz_boundary_jacobian(const nz_problem *p, const double *U, double *bjac, void *stream);

/* dt_e = factor * 2 A_e / max(wave_e, 1e-10) (src/timestepping/step.py:171-183). */
int nz_time_step(const nz_problem *p, const double *wave, double factor, double *dt,
                 void *stream);

/* ---- mesh connectivity and geometry on the device -------------------------
 * (src/geometry/mesh.py:77-208 build_edges; src/spatial/operators.py:209-342
 * build_operators).  Sizes that are only known after the build (interior and
 * boundary edge counts) come back in `counts`; outputs are sized for the
 * worst case by the caller.  Errors: nz_mesh_last_error(). */
#define NZ_TAG_INFLOW 1
#define NZ_TAG_OUTFLOW 2
#define NZ_TAG_WALL 3
#define NZ_TAG_AXIS 4

typedef struct {
    int64_t n_elem;
    int64_t n_nodes;
    int32_t nface;            /* 3 triangles, 4 quadrilaterals */
    int32_t nodes_per_elem;   /* geometry nodes per element */
    const int64_t *elem_nodes;  /* (E, nodes_per_elem) */
    int32_t face_ends[8];     /* corners of face f: local nodes face_ends[2f], face_ends[2f+1] */
    const double *tag_coords; /* (n_nodes, 2): the coordinates the boundary tagger classifies */
    double tag_length, tag_tol;  /* nozzle_boundary_tagger (src/geometry/mesh.py:340-358) */
} nz_mesh_desc;

typedef struct {
    /* the reference's EdgeConnectivity (src/geometry/mesh.py:54-74), int64;
     * interior arrays sized (E*nface/2, 2), boundary arrays (E*nface,) */
    int64_t *iedge_elem, *iedge_face;
    int64_t *bedge_elem, *bedge_face, *bedge_tag;
    int64_t *face_edge, *face_side, *face_neighbour;  /* (E, nface) */
    /* owner element and local face per edge, interior then boundary (E*nface,);
     * nullable together */
    int64_t *edge_elem, *edge_face;
    /* the element kernel's maps (nz_problem fe_nbr, fe_info, bedge); nullable together */
    int32_t *fe_nbr, *fe_info, *bedge;
    /* (8,): n_interior, n_boundary, edges with tag 1, 2, 3, 4, faces shared
     * by more than two elements (non-zero: the mesh is not conforming) */
    int64_t *counts;
} nz_edge_maps;

int64_t nz_build_edges_workspace_bytes(int64_t n_elem, int32_t nface);

/* Edge lists and face maps of a conforming mesh, bit-identical to
 * build_edges (src/geometry/mesh.py:77-208): faces matched on their corner
 * pair by a stable sort of the keys, interior edges in key order owned by
 * the lower element, boundary faces in key order grouped by tag. */
int nz_build_edges(const nz_mesh_desc *m, const nz_edge_maps *out, void *work,
                   int64_t work_bytes, void *stream);

/* Offsets (doubles) of the reference geometry tables inside
 * nz_geometry_desc.tables (src/spatial/operators.py:46-135). */
typedef struct {
    int32_t dgeom_dxi_vol;      /* (NQ, nqv) */
    int32_t dgeom_deta_vol;     /* (NQ, nqv) */
    int32_t geom_vol;           /* (NQ, nqv) */
    int32_t w_vol;              /* (nqv,) */
    int32_t phi_vol;            /* (N, nqv) */
    int32_t dgeom_dsigma_face;  /* (nface, NQ, nqf) */
    int32_t geom_face;          /* (nface, NQ, nqf) */
    int32_t w_face;             /* (nqf,) */
} nz_geometry_offsets;

typedef struct {
    int64_t n_elem, n_edge;
    int32_t nq, nqv, nqf, nbf, nface;
    int32_t affine, straight;   /* as in nz_problem */
    int32_t pad0;
    const int64_t *elem_nodes;  /* (E, nq) */
    const double *coords;       /* (n_nodes, 2) */
    const int64_t *edge_elem, *edge_face;  /* owner element and face per edge (nz_edge_maps) */
    const double *tables;
    nz_geometry_offsets off;
} nz_geometry_desc;

typedef struct {
    /* the element kernel's layout (nz_problem): geo, area, minv and mean_w
     * (non-affine only), edge_geo */
    double *geo, *area, *minv, *mean_w, *edge_geo;
    /* the reference's Operators layout, each nullable:
     * det (E,nqv), metrics (E,nqv,4) = xi_x, xi_y, eta_x, eta_y, wdet (E,nqv),
     * xy_vol (E,nqv,2), mean_w_full (E,N) (also for affine elements),
     * normal (F,nqf,2), jac (F,nqf), length (F,), xy_face (F,nqf,2) */
    double *det, *metrics, *wdet, *xy_vol, *mean_w_full;
    double *normal, *jac, *length, *xy_face;
    int64_t *bad_jacobian;      /* (1,): volume points with det <= 0 */
} nz_geometry_out;

/* Per-element and per-edge geometry from node coordinates alone: the
 * build_operators quantities (src/spatial/operators.py:245-318) in the
 * element kernel's layout, the inverse mass matrices of non-affine elements
 * by Cholesky.  Reused topology: call again with new coordinates. */
int nz_build_geometry(const nz_geometry_desc *g, const nz_geometry_out *out, void *stream);

/* Last error of nz_build_edges / nz_build_geometry on this thread. */
const char *nz_mesh_last_error(void);

/* ---- observables (src/post/postprocess.py:174-255, src/sensitivity.py:96-130) */
#define NZ_OBJ_THRUST 0
#define NZ_OBJ_THRUST_COEFFICIENT 1
#define NZ_OBJ_MASS_FLOW 2
#define NZ_OBJ_EXIT_MACH 3
#define NZ_OBJ_EXIT_PRESSURE 4

/* Workspace (doubles) of nz_performance_integrals and nz_objective. */
int64_t nz_observables_work_doubles(const nz_problem *p);

/* out[0] = integral over the domain of (p / (rho^gamma s_t) - 1)^2 (the
 * entropy error times the area, entropy_error, postprocess.py:174-190);
 * out[1] = integral over the outflow of p / rho ds (the exit temperature
 * ratio's numerator times R).  Summed in a fixed order. */
int nz_performance_integrals(const nz_problem *p, const double *U, double s_t, double *work,
                             double *out, void *stream);

/* J (one double), dJ/dU (E, N, 4) and dJ/d(n_x, n_y, |dX/dsigma|) (F, nqf, 3)
 * of an objective (objective_value, src/sensitivity.py:96-130; ids
 * NZ_OBJ_*): thrust 2 sum_{out,in} int (rho v_x v_n + p n_x) ds, its
 * coefficient (scale = 1 / (p_t A_throat)), mass flow -2 int_in rho v_n ds,
 * and the outflow averages of the Mach number and of p (scale = 1 / p_t).
 * The density and pressure floors contribute no derivative when active.
 * Interior edges get zero geometry derivatives. */
int nz_objective(const nz_problem *p, const double *U, int32_t objective, double scale,
                 double *work, double *J, double *dU, double *dG, void *stream);
const char *nz_observables_last_error(void);

/* ---- quasi-1D initial state on the device -------------------------------- */
#define NZ_CONTOUR_HERMITE 0   /* bell, smooth, moc */
#define NZ_CONTOUR_CONICAL 1
#define NZ_CONTOUR_BEZIER 2    /* cubic with two weights */
#define NZ_CONTOUR_BEZIER9 3   /* degree-9 Bernstein diverging wall */
#define NZ_CONTOUR_ANALYTIC 4

typedef struct {
    int32_t contour;            /* NZ_CONTOUR_* */
    int32_t n_points;           /* quasi-1D stations */
    const double *x, *density, *velocity, *pressure;  /* (n_points,), x increasing */
    /* wall (src/geometry/contour.py:264-311), parameters as the host computes them */
    double length, throat_x, y_throat, y_inlet, y_exit;
    double slope_throat, slope_exit;  /* Hermite end slopes: tan(theta) (1 - throat_x) length */
    double bezier_c1, bezier_c2;      /* cubic Bezier interior control heights */
    double ctrl9[10];                 /* degree-9 control heights, throat to exit */
    double binom9[10];                /* C(9, k) */
    double pi, e2m1;                  /* pi and exp(2) - 1 for the analytic wall */
    double gamma;
    double h;                         /* central-difference step of the wall slope, 1e-6 L */
} nz_quasi1d_desc;

/* Conserved point values (n, 4) of the quasi-1D initial state at the points
 * xy (n, 2): the step before nz_project in quasi1d_initial
 * (src/timestepping/initialize.py:68-102).  Hermite and conical walls are
 * evaluated in the host's operation order without FP contraction
 * (bit-identical values); Bezier walls round integer powers correctly. */
int nz_quasi1d_values(const nz_quasi1d_desc *d, const double *xy, int64_t n, double *values,
                      void *stream);
const char *nz_initial_last_error(void);

/* Last error message of this thread (static storage; never NULL). */
const char *nz_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* NOZZLE_HIP_H */
