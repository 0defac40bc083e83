"""The ``hip`` backend: the reference ``Backend`` interface on MI355X kernels.

``HipBackend(ops, flow, opts)`` keeps the constructor, attributes and methods
of the reference's backend ABC (``src/backends/base.py:81-201``) and of its
compiled CPU backend (``src/backends/numba_backend.py:50-337``):

* primitives ``residual(U) -> (R, wave)``, ``inverse_mass(R)``, ``limit(U)``,
  ``time_step(wave)``, ``norm(A)``, ``asarray``, ``to_numpy``;
* integrators ``rate``, ``euler_step``, ``ssprk3_step``, ``make_step`` and
  ``run(U, n_steps, scheme) -> (U, res_last)``;
* state ``name``, ``scheme``, ``n_limited``, ``n_mean_repaired``,
  ``n_limit_calls``, ``limiter_activity()``, ``component_weight``.

Arrays go in and come out in the same kind: NumPy in, NumPy out (so the
reference's march and tests work unchanged), or ``torch`` device tensors in,
device tensors out (no host round trip).  ``run`` copies the state in once and
out once and never aliases its own buffers in what it returns
(``numba_backend.py:283-303``).

Every numerical call goes through ``libnozzle_hip.so``; there is no CPU path.
An unknown flux raises ``NotImplementedError`` and an unknown scheme
``ValueError``, as in the reference.
"""

from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _lib
from .config import (SCHEMES, component_weights, recommended_scheme, reference_floors,
                     step_coefficient)
from .operators import element_metrics

FLUX_IDS = {"roe": 0, "hllc": 1, "ausmup": 2}


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the hip backend needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _tables(rd) -> tuple[np.ndarray, dict]:
    """Reference tables, concatenated; offsets in doubles."""
    phi, dxi, deta, w = rd.phi_vol, rd.dphi_dxi_vol, rd.dphi_deta_vol, rd.w_vol
    mass_ref = (phi * w[None, :]) @ phi.T
    minv_ref = np.linalg.inv(mass_ref)
    parts = {
        "phi_vol": phi,
        "dxi": dxi,
        "deta": deta,
        "axi": minv_ref @ (dxi * w[None, :]),
        "aeta": minv_ref @ (deta * w[None, :]),
        "phi_face": rd.phi_face,
        "lift": np.einsum("ij,sfjq->sfiq", minv_ref, rd.phi_face),
        "w_vol": w,
        "w_face": rd.w_face,
        "mean_w": (phi @ w) / w.sum(),
        "minv_ref": minv_ref,
    }
    # point-major copies of the volume tables: one contiguous block per point
    parts["vol_t"] = np.stack([phi, dxi, deta, parts["axi"], parts["aeta"]], axis=0).transpose(2, 0, 1)
    off, chunks, pos = {}, [], 0
    for k, v in parts.items():
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        off[k] = pos
        chunks.append(v)
        pos += v.size
    return np.concatenate(chunks), off


TILE = 16                      # elements per tile of the element kernel
XCDS = 8                       # L2 domains a launch is dealt round-robin across
LOCALITY_MIN_ELEMENTS = 1 << 14
# (element kind, geometry order, order) -> the residual-VJP tile shape the
# autotuner measured fastest with the Roe flux in this process (0 one-wave
# tiles of 16 elements, 1 four-wave tiles); see HipBackend.tune_vjp
_VJP_KERNEL: dict = {}
# (element kind, order, flux, structured grid) -> True when the element
# kernel on patch tiles measured faster than on tiles (affine meshes of at
# least LOCALITY_MIN_ELEMENTS elements); see HipBackend.tune_element_kernel
_DG_PATCHES: dict = {}
# the same key -> milliseconds of the three timed steps, (tiles, patches)
_DG_COSTS: dict = {}


def _morton2(ix: np.ndarray, iy: np.ndarray) -> np.ndarray:
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0xFFFF)
        for sh, m in ((8, 0x00FF00FF), (4, 0x0F0F0F0F), (2, 0x33333333), (1, 0x55555555)):
            v = (v | (v << np.uint64(sh))) & np.uint64(m)
        return v
    return spread(ix) | (spread(iy) << np.uint64(1))


def locality_tiles(coords: np.ndarray, elem_nodes: np.ndarray, E: int,
                   n_domains: int = XCDS) -> np.ndarray:
    """Tile table ``(T, 3)``: ``(first element, end, norm slot)`` per workgroup.

    Tiles are the consecutive 16-element blocks (slot = block index).  They
    are ordered along a Morton curve of their centroids, and the order is dealt
    so that workgroup ``b``, which runs on L2 domain ``b % n_domains``, takes
    the ``b // n_domains``-th tile of that domain's contiguous share.
    """
    T = (E + TILE - 1) // TILE
    starts = np.arange(0, E, TILE)
    cent = coords[elem_nodes].mean(axis=1)
    cnt = np.diff(np.append(starts, E))
    tc = np.stack([np.add.reduceat(cent[:, k], starts) / cnt for k in range(2)], axis=1)
    lo, span = tc.min(axis=0), np.ptp(tc, axis=0) + 1e-300
    q = np.minimum((tc - lo) / span * 65536.0, 65535.0).astype(np.int64)
    order = np.argsort(_morton2(q[:, 0], q[:, 1]), kind="stable")
    share = -(-T // n_domains)
    b = np.arange(T)
    pos = (b % n_domains) * share + b // n_domains
    ok = pos < T
    launch = np.empty(T, dtype=np.int64)
    launch[ok] = order[pos[ok]]
    taken = np.zeros(T, dtype=bool)
    taken[pos[ok]] = True
    launch[~ok] = order[~taken]
    e0 = launch * TILE
    return np.stack([e0, np.minimum(e0 + TILE, E), launch], axis=1).astype(np.int64)


PATCH_INTS = 18  # per patch: 16 element ids (-1 padded), count, norm slot


def patch_order(topo, coords: np.ndarray, E: int) -> np.ndarray:
    """Element order whose consecutive runs of 16 form compact patches.

    On the structured triangle grids (``topo.logical`` names ``nx_cells`` by
    ``nr_cells`` cells, elements ``c`` and ``c + ncx ncy`` being the lower and
    upper triangle of cell ``c = i ncy + j``) the cells follow a Morton curve
    of ``(j, i)`` with both triangles of a cell together, so aligned runs of
    16 are blocks of 2 x 4 cells whose four cells along ``j`` have consecutive
    ids (rows of U, dt and flags in runs of four elements); otherwise the
    Morton curve of the element centroids.  The order only decides which
    elements share a workgroup.
    """
    lg = getattr(topo, "logical", None) or {}
    ncx, ncy = int(lg.get("nx_cells", 0)), int(lg.get("nr_cells", 0))
    e = np.arange(E, dtype=np.int64)
    if topo.kind == "tri" and ncx > 0 and ncy > 0 and E == 2 * ncx * ncy and max(ncx, ncy) <= 65536:
        c = e % (ncx * ncy)
        key = (_morton2(c % ncy, c // ncy) << np.uint64(1)) | (e // (ncx * ncy)).astype(np.uint64)
    else:
        cent = coords[topo.elem_nodes].mean(axis=1)
        lo, span = cent.min(axis=0), np.ptp(cent, axis=0) + 1e-300
        qc = np.minimum((cent - lo) / span * 65536.0, 65535.0).astype(np.int64)
        key = _morton2(qc[:, 0], qc[:, 1])
    return np.argsort(key, kind="stable")


def patch_table(order: np.ndarray, n_domains: int = XCDS) -> np.ndarray:
    """Patch table ``(P, 18)`` int32 for the element kernel: the elements of
    ``order`` in runs of 16 (ids, -1 padding, count, norm slot = run index),
    dealt across the L2 domains as :func:`locality_tiles` deals tiles."""
    E = order.shape[0]
    n = (E + TILE - 1) // TILE
    ids = np.full(n * TILE, -1, dtype=np.int64)
    ids[:E] = order
    ids = ids.reshape(n, TILE)
    cnt = np.full(n, TILE, dtype=np.int64)
    cnt[-1] = E - TILE * (n - 1)
    share = -(-n // n_domains)
    b = np.arange(n)
    pos = (b % n_domains) * share + b // n_domains
    ok = pos < n
    launch = np.empty(n, dtype=np.int64)
    launch[ok] = pos[ok]
    taken = np.zeros(n, dtype=bool)
    taken[pos[ok]] = True
    launch[~ok] = np.flatnonzero(~taken)
    return np.concatenate([ids[launch], cnt[launch, None], launch[:, None]], axis=1).astype(np.int32)


def _least_squares_hessenberg(H: np.ndarray, beta: float) -> tuple[np.ndarray, float]:
    """``min_y || beta e1 - H y ||`` for an upper Hessenberg ``H`` of shape
    ``(k + 1, k)`` by Givens rotations; returns ``(y, residual norm)``.  A
    zero subdiagonal entry (an exact Krylov space) ends the columns there."""
    k = H.shape[1]
    sub = np.flatnonzero(np.diag(H, -1) == 0.0)
    if sub.size:
        k = int(sub[0]) + 1
    R = H[:k + 1, :k].copy()
    g = np.zeros(k + 1)
    g[0] = beta
    for j in range(k):
        a, b = R[j, j], R[j + 1, j]
        den = math.hypot(a, b)
        c, sn = (1.0, 0.0) if den == 0.0 else (a / den, b / den)
        rj, rj1 = R[j, j:].copy(), R[j + 1, j:].copy()
        R[j, j:] = c * rj + sn * rj1
        R[j + 1, j:] = -sn * rj + c * rj1
        g[j], g[j + 1] = c * g[j], -sn * g[j]
    y = np.zeros(k)
    for i in range(k - 1, -1, -1):  # back substitution
        y[i] = (g[i] - R[i, i + 1:k] @ y[i + 1:k]) / R[i, i]
    return y, abs(g[k])


# problems with fewer element-basis pairs than this chain their stages'
# boundary fluxes (HipBackend.chain_bflux)
CHAIN_MAX_DOFS = 1 << 22


class HipBackend:
    """MI355X execution of the DG residual and the pseudo-time march."""

    name = "hip"

    def __init__(self, ops, flow, opts, *, device: int | None = None):
        torch = _torch()
        self._torch = torch
        self.lib = _lib.load()
        self.ops, self.flow, self.opts = ops, flow, opts
        self.n_limited = 0
        self.n_mean_repaired = 0
        self.n_limit_calls = 0
        self.component_weight = np.asarray(component_weights(flow)).reshape(1, 1, 4)
        rd = ops.ref
        self.scheme = recommended_scheme(rd.order) if opts.scheme is None else opts.scheme
        flux = getattr(flow, "flux", "roe")
        if flux not in FLUX_IDS:
            raise NotImplementedError(
                f"the hip kernels implement {sorted(FLUX_IDS)}, so flux={flux!r} would silently "
                "give a different answer here; add a kernel for it"
            )
        if opts.limiter not in ("none", "positivity", "superbee"):
            raise ValueError(f"unknown limiter {opts.limiter!r}")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        f64 = dict(dtype=torch.float64, device=self.device)
        topo = ops.topology
        kind, q_geo = topo.kind, topo.geometry_order
        E, N, nqv, nqf, nface = ops.n_elem, rd.n_basis, rd.n_qvol, rd.n_qface, topo.n_faces
        self.E, self.N = E, N
        self.affine = kind == "tri" and q_geo == 1
        self.straight = q_geo == 1
        ni = ops.n_interior
        F = ni + ops.n_boundary
        if F >= 2**28 or E >= 2**31:
            raise ValueError("mesh too large for the 32-bit face maps")

        dev_arrays = getattr(ops, "device_arrays", None)
        if dev_arrays is not None:
            # operators built on the device (device_mesh.DeviceOperators):
            # every array is already in the element kernel's layout
            self._init_device_geometry(ops, dev_arrays, nqf)
        else:
            self._init_host_geometry(ops, rd, topo, topo.edges, E, N, nqv, nqf, nface, ni)
        tables, off = _tables(rd)
        self._keep["tables"] = torch.from_numpy(np.ascontiguousarray(tables)).to(self.device)
        self._keep["bflux"] = torch.empty(max(ops.n_boundary, 1) * nqf * 5, **f64)
        self.n_boundary_points = ops.n_boundary * nqf

        def ptr(k):
            t = self._keep[k]
            return C.c_void_p(t.data_ptr() if t is not None else 0)

        at2 = flow.gamma * flow.Rgas * flow.total_temperature
        rho_floor, p_floor = reference_floors(flow)
        w = component_weights(flow)
        P = _lib.Problem()
        P.n_elem, P.n_edge = E, F
        P.nface, P.nbf, P.nqv, P.nqf = nface, N, nqv, nqf
        P.affine, P.straight = int(self.affine), int(self.straight)
        for k in ("geo", "area", "minv", "mean_w", "edge_geo", "fe_nbr", "fe_info", "bedge",
                  "bflux", "tables"):
            setattr(P, k, ptr(k))
        P.n_int = ni
        for k, v in off.items():
            setattr(P.off, k, v)
        P.gamma, P.efix = flow.gamma, flow.entropy_fix
        P.low_mach = getattr(flow, "hllc_low_mach", 0.0)
        P.flux_id = FLUX_IDS[flux]
        self.limiting = opts.limiter != "none" and rd.order > 0
        self.superbee = self.limiting and opts.limiter == "superbee"
        P.limiter = (2 if self.superbee else 1) if self.limiting else 0
        # Consecutive stages hand the boundary fluxes of their output to the
        # next instead of launching the boundary pre-pass (bit-identical).  It
        # pays where the pre-pass is a sizeable share of a stage (a 4,096-
        # element P1 channel: 7 % per step); on the 2^20-element P3 channel
        # the heavier stage epilogue costs more than the 5 us pre-pass it
        # removes (0.5 % per step), so large problems keep the pre-pass.
        self.chain_bflux = E * N < CHAIN_MAX_DOFS
        self.tvb = float(getattr(opts, "tvb_constant", 50.0))
        P.rho_t_bf = float(flow.stagnation_density) if getattr(flow, "backflow", False) else 0.0
        P.p_t_bf, P.band_bf = flow.total_pressure, 0.05
        P.at2, P.at = at2, math.sqrt(at2)
        P.rho_t = flow.gamma * flow.total_pressure / at2
        P.ca, P.sa = math.cos(flow.inflow_angle), math.sin(flow.inflow_angle)
        P.p_back = flow.back_pressure_ratio * flow.total_pressure
        P.rho_floor, P.p_floor, P.fraction = rho_floor, p_floor, 1e-8
        # Lebesgue constant of the basis over the probe set (numba_backend.py:115-120)
        P.lebesgue = float(max(np.abs(rd.phi_vol).sum(axis=0).max(),
                               np.abs(rd.phi_face).sum(axis=2).max()))
        P.bisect_steps = 12
        for s in range(4):
            P.wnorm[s] = w[s]
        self._vjp_key = (topo.kind, int(getattr(topo, "geometry_order", 1)), int(rd.order),
                         int(P.flux_id))
        P.vjp_kernel = _VJP_KERNEL.get(self._vjp_key, 0)
        self.P = P
        self._pp = C.byref(P)

        # -- working set, allocated once (numba_backend.py:122-134)
        shape = (E, N, 4)
        self._S = [torch.empty(shape, **f64) for _ in range(4)]
        self._F = torch.empty(shape, **f64)
        self._wave = torch.empty(E, **f64)
        self._dt = torch.empty(E, **f64)
        nscr = int(self.lib.nz_scratch_doubles(self._pp))
        if nscr < 0:
            _lib.check(-1, "nz_scratch_doubles")
        self._scratch = torch.empty(nscr, **f64)
        self._res = torch.empty(1, **f64)
        # positivity scaled, cell means repaired, superbee slopes scaled
        self._counters = torch.zeros(3, dtype=torch.int64, device=self.device)
        self._sb_work = None
        if self.superbee:
            nsb = int(self.lib.nz_superbee_work_doubles(self._pp))
            self._sb_work = torch.empty(nsb, **f64)
        self._flags = torch.zeros(E, dtype=torch.uint8, device=self.device)
        # captured march chunks (see _advance_chunk); use_graphs=False runs
        # every chunk eagerly
        self.use_graphs = True
        self._graphs: dict = {}
        self._graph_seen: set = set()
        # launch order of the element kernel's tiles: spatially coherent, and
        # one contiguous run of it per XCD, so that the neighbour columns a
        # tile gathers were just read through the same L2 (norm slots keep the
        # element order, so results do not depend on the launch order)
        self._tiles = None
        self._patches = None
        if E >= LOCALITY_MIN_ELEMENTS:
            tab = locality_tiles(np.asarray(ops.node_coords), np.asarray(topo.elem_nodes), E)
            self._tiles = torch.from_numpy(tab.reshape(-1)).to(self.device)
            P.tiles, P.n_tiles = C.c_void_p(self._tiles.data_ptr()), tab.shape[0]
            P.tiles_cover_mesh = 1  # a permutation of every tile: the adjoint uses it too
            if self.affine:
                self.set_patches(_DG_PATCHES.get(self._dg_key, False))

    def set_patches(self, on: bool) -> None:
        """Run the element kernel on patch tiles (affine meshes: each shared
        edge's flux solved once) or on consecutive / locality-ordered tiles.
        The two give bit-identical states; the residual norm's partial sums are
        grouped by workgroup, so it agrees to rounding.  Drops captured march
        chunks, whose launches hold the old problem."""
        if on and not self.affine:
            raise ValueError("patch tiles need an affine (straight-sided triangle) mesh")
        if on and self._patches is None:
            ops = self.ops
            order = patch_order(ops.topology, np.asarray(ops.node_coords), self.E)
            self._patches = self._torch.from_numpy(patch_table(order).reshape(-1)).to(self.device)
        self.P.patches = C.c_void_p(self._patches.data_ptr() if on else 0)
        self.P.n_patches = self._patches.numel() // PATCH_INTS if on else 0
        nscr = int(self.lib.nz_scratch_doubles(self._pp))  # one norm partial per patch
        if nscr < 0:
            _lib.check(-1, "nz_scratch_doubles")
        if self._scratch.numel() < nscr:
            self._scratch = self._torch.empty(nscr, dtype=self._torch.float64, device=self.device)
        self._graphs.clear()
        self._graph_seen.clear()

    @property
    def patches(self) -> bool:
        return bool(self.P.patches)

    @property
    def _dg_key(self):
        lg = getattr(self.ops.topology, "logical", None) or {}
        return (self.ops.topology.kind, int(self.ops.ref.order), int(self.P.flux_id),
                bool(lg.get("nx_cells")))

    def tune_element_kernel(self, U, *, warm_s: float = 0.25, rounds: int = 5,
                            steps: int = 3) -> bool:
        """Autotuner for the element kernel's work shape on affine meshes of
        at least LOCALITY_MIN_ELEMENTS elements: after ``warm_s`` seconds of
        alternating launches (the device's clocks settle over the first
        hundreds of launches), ``rounds`` alternating timings of ``steps``
        fused SSP-RK3 steps from ``U`` on locality-ordered tiles and on patch
        tiles; the shape with the smaller median is kept for this (element
        kind, order, flux, structured or not) for the rest of the process.
        Returns True for patches.  The two give the same states bit for bit.
        Smaller or curved meshes keep tiles."""
        if not self.affine or self._tiles is None:
            return False
        key = self._dg_key
        if key not in _DG_PATCHES:
            torch = self._torch
            self._check_shape(U, "state", "U")
            u, _ = self._in(U)
            st = torch.cuda.current_stream(self.device)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

            def timed(on: bool) -> float:
                self.set_patches(on)
                self._S[0].copy_(u)
                e0.record(st)
                self.advance(self._S[0], self._S[1], steps, "ssprk3")
                e1.record(st)
                e1.synchronize()
                return e0.elapsed_time(e1)

            t_end = time.perf_counter() + warm_s
            while time.perf_counter() < t_end:
                timed(False)
                timed(True)
            ta, tb = [], []
            for _ in range(max(int(rounds), 1)):
                ta.append(timed(False))
                tb.append(timed(True))
            cost = (float(np.median(ta)), float(np.median(tb)))
            _DG_PATCHES[key] = cost[1] < cost[0]
            _DG_COSTS[key] = cost
        self.set_patches(_DG_PATCHES[key])
        return _DG_PATCHES[key]

    # ------------------------------------------------------------ set-up
    def _init_host_geometry(self, ops, rd, topo, ed, E, N, nqv, nqf, nface, ni):
        """The element kernel's compact geometry and maps from host operators."""
        torch = self._torch
        X = np.asarray(ops.node_coords, dtype=np.float64)
        det, xix, xiy, etx, ety = element_metrics(rd, X[topo.elem_nodes])
        if self.affine:
            geo = np.stack([xix[:, 0], xiy[:, 0], etx[:, 0], ety[:, 0], det[:, 0]], axis=-1)
        else:
            geo = np.stack([xix, xiy, etx, ety, rd.w_vol[None, :] * det], axis=-1)
        minv = mean_w = None
        if not self.affine:
            minv = getattr(ops, "inv_mass", None)
            if minv is None:
                wdet = rd.w_vol[None, :] * det
                minv = np.linalg.inv(np.einsum("iq,jq,eq->eij", rd.phi_vol, rd.phi_vol, wdet))
            # the device keeps the upper triangles, row-major (i <= j)
            iu = np.triu_indices(rd.n_basis)
            minv = np.asarray(minv)[:, iu[0], iu[1]]
            mean_w = np.asarray(ops.mean_weights)
        nrm = np.asarray(ops.edge_normal)
        jac = np.asarray(ops.edge_jac)
        if self.straight:
            edge_geo = np.stack([nrm[:, 0, 0], nrm[:, 0, 1], jac[:, 0]], axis=-1)
        else:
            edge_geo = np.stack([nrm[..., 0], nrm[..., 1], jac], axis=-1)

        # -- maps: neighbour (or -tag) and (edge << 3 | side << 2 | other face)
        fe, fs = np.asarray(ed.face_edge), np.asarray(ed.face_side)
        nbr = np.asarray(ed.face_neighbour).astype(np.int64).copy()
        bnd = fe >= ni
        nbr[bnd] = -np.asarray(ed.bedge_tag)[fe[bnd] - ni]
        oface = np.zeros_like(fe)
        inner = ~bnd
        iface = np.asarray(ed.iedge_face)
        oface[inner] = np.where(fs[inner] == 0, iface[fe[inner], 1], iface[fe[inner], 0])
        info = (fe << 3) | (fs << 2) | oface
        bedge = np.stack([np.asarray(ed.bedge_elem),
                          (np.asarray(ed.bedge_face) << 3) | np.asarray(ed.bedge_tag)], axis=-1)

        def dev(a, dtype=None):
            if a is None:
                return None
            t = torch.from_numpy(np.ascontiguousarray(a))
            return t.to(device=self.device, dtype=dtype or t.dtype)

        self._keep = {
            "geo": dev(geo, torch.float64),
            "area": dev(np.asarray(ops.elem_area), torch.float64),
            "minv": dev(minv, torch.float64),
            "mean_w": dev(mean_w, torch.float64),
            "edge_geo": dev(edge_geo, torch.float64),
            "fe_nbr": dev(nbr.astype(np.int32)),
            "fe_info": dev(info.astype(np.int32)),
            "bedge": dev(bedge.astype(np.int32)),
        }
        self.area_total = float(np.sum(np.asarray(ops.elem_area)))

    def _init_device_geometry(self, ops, arrays, nqf):
        """Geometry and maps already on the device (``DeviceOperators``)."""
        keep = {}
        for k in ("geo", "area", "minv", "mean_w", "edge_geo", "fe_nbr", "fe_info", "bedge"):
            t = arrays.get(k)
            if t is not None and t.device != self.device:
                raise ValueError(f"device operator array {k} is on {t.device}, not {self.device}")
            keep[k] = t
        self._keep = keep
        self.area_total = float(ops.elem_area_total)

    # ------------------------------------------------------------ helpers
    def _stream(self):
        return C.c_void_p(self._torch.cuda.current_stream(self.device).cuda_stream)

    def _in(self, A, shape=None, what: str = "array"):
        """(device tensor, was_numpy).

        ``shape``: the shape the kernels assume, ``"state"`` for ``(E, N, 4)``
        or ``"elem"`` for ``(E,)``.  The C ABI takes no sizes, so this is the
        guard that keeps an array of another mesh or order from being read or
        written past its end on the device; it raises before any transfer.
        """
        if shape is not None:
            self._check_shape(A, shape, what)
        torch = self._torch
        if isinstance(A, torch.Tensor):
            t = A.to(device=self.device, dtype=torch.float64)
            return t.contiguous(), False
        return torch.from_numpy(np.ascontiguousarray(A, dtype=np.float64)).to(self.device), True

    def _check_shape(self, A, shape, what: str) -> None:
        want = {"state": (self.E, self.N, 4), "elem": (self.E,)}.get(shape, shape)
        got = tuple(np.shape(A))
        if got != tuple(want):
            raise ValueError(f"{what} must have shape {tuple(want)} for this mesh and order, "
                             f"got {got}")

    def _out(self, t, as_numpy: bool):
        return t.cpu().numpy() if as_numpy else t

    @staticmethod
    def _p(t) -> C.c_void_p:
        return C.c_void_p(t.data_ptr())

    def _chk(self, rc, what):
        _lib.check(rc, what)

    # ------------------------------------------------------------ state
    def limiter_activity(self) -> float:
        if self.n_limit_calls == 0:
            return float("nan")
        return self.n_limited / self.n_limit_calls

    def positivity_scale(self) -> float:
        """``base.py:123-134``: extra dt factor while a limiter is on."""
        pc = getattr(self.opts, "positivity_cfl", 0.0)
        if pc > 0.0 and self.opts.limiter != "none" and self.ops.ref.order > 0:
            return float(pc)
        return 1.0

    def dt_factor(self, scheme: str | None = None) -> float:
        sch = self.scheme if scheme is None else scheme
        return self.positivity_scale() * step_coefficient(self.ops.ref.order, self.opts.cfl, sch)

    def asarray(self, A):
        return self._in(A)[0]

    def to_numpy(self, A) -> np.ndarray:
        torch = self._torch
        if isinstance(A, torch.Tensor):
            return A.detach().cpu().numpy()
        return np.asarray(A)

    # ------------------------------------------------------------ primitives
    def residual(self, U):
        """``(R, wave)`` with ``M dU/dt = -R`` (``assembly.py:153-157``)."""
        u, npy = self._in(U, "state", "U")
        R = self._torch.empty_like(u)
        wave = self._torch.empty(self.E, dtype=u.dtype, device=self.device)
        self._chk(self.lib.nz_residual(self._pp, self._p(u), self._p(R), self._p(wave),
                                       self._stream()), "nz_residual")
        return self._out(R, npy), self._out(wave, npy)

    def rate(self, U):
        """``-M^-1 R`` and the wave-speed sum, in one pass (``element_rate``)."""
        u, npy = self._in(U, "state", "U")
        Fr = self._torch.empty_like(u)
        wave = self._torch.empty(self.E, dtype=u.dtype, device=self.device)
        self._chk(self.lib.nz_rate(self._pp, self._p(u), self._p(Fr), self._p(wave),
                                   self._stream()), "nz_rate")
        return self._out(Fr, npy), self._out(wave, npy)

    def inverse_mass(self, R):
        r, npy = self._in(R, "state", "R")
        out = self._torch.empty_like(r)
        self._chk(self.lib.nz_inverse_mass(self._pp, self._p(r), self._p(out), self._stream()),
                  "nz_inverse_mass")
        return self._out(out, npy)

    def project(self, values):
        """L2 projection of point values ``(E, nqv, 4)`` onto the basis, on the
        device (``project``, ``src/timestepping/initialize.py:41-46``)."""
        rd = self.ops.ref
        v, npy = self._in(values, (self.E, rd.n_qvol, 4), "values")
        out = self._torch.empty((self.E, rd.n_basis, 4), dtype=self._torch.float64,
                                device=self.device)
        self._chk(self.lib.nz_project(self._pp, self._p(v), self._p(out), self._stream()),
                  "nz_project")
        return self._out(out, npy)

    def performance(self, U, geom=None, quasi1d=None):
        """Every performance field of ``U`` (an ``observables.Performance``,
        ``src/post/postprocess.py:193-255``) from integrals on the device:
        the boundary integrals, the entropy integral over the domain and the
        exit temperature integral.  ``geom`` (the ``NozzleGeometry``) gives
        the throat-based coefficients and the quasi-1D ideal thrust."""
        from .observables import assemble

        torch = self._torch
        u, _ = self._in(U, "state", "U")
        f64 = dict(dtype=torch.float64, device=self.device)
        bnd = torch.empty(8, **f64)
        vol = torch.empty(2, **f64)
        work = torch.empty(max(int(self.lib.nz_observables_work_doubles(self._pp)), 1), **f64)
        st = self._stream()
        flow = self.flow
        g = flow.gamma
        s_t = flow.total_pressure ** (1.0 - g) * (flow.Rgas * flow.total_temperature) ** g
        self._chk(self.lib.nz_boundary_integrals(self._pp, self._p(u), self._p(bnd), st),
                  "nz_boundary_integrals")
        rc = self.lib.nz_performance_integrals(self._pp, self._p(u), s_t, self._p(work),
                                               self._p(vol), st)
        if rc != 0:
            raise _lib.HipError(f"nz_performance_integrals failed ({rc}): "
                                f"{self.lib.nz_observables_last_error().decode(errors='replace')}")
        mi, mo, wall, qi, qo, area_e, mach, p = bnd.cpu().tolist()
        ent, pr = vol.cpu().tolist()
        return assemble(flow, thrust=2.0 * (mo + mi), thrust_wall=2.0 * wall, mdot_in=2.0 * -qi,
                        mdot_out=2.0 * qo, mach_int=mach, p_int=p, t_int=pr / flow.Rgas,
                        area_e=area_e, entropy_int=ent, area=self.area_total, geom=geom,
                        quasi1d=quasi1d)

    def objective(self, U, name: str = "thrust", *, geom=None):
        """``(J, dJ/dU, dJ/d(edge geometry))`` of a named objective on the
        device (``nz_objective``; ``objective_value``,
        ``src/sensitivity.py:96-130``): ``J`` a float, ``dJ/dU`` a device
        tensor ``(E, N, 4)``, the edge cotangent ``(F, nqf, 3)`` a device
        tensor.  ``thrust_coefficient`` needs ``geom`` (the throat height)."""
        from .sensitivity import OBJECTIVE_IDS, OBJECTIVES

        if name not in OBJECTIVES:
            raise ValueError(f"unknown objective {name!r}; choose from {OBJECTIVES}")
        torch = self._torch
        u, _ = self._in(U, "state", "U")
        oid = OBJECTIVE_IDS[name]
        scale = 1.0
        if name == "thrust_coefficient":
            if geom is None:
                raise ValueError("the thrust coefficient needs the geometry (throat height)")
            scale = 1.0 / (self.flow.total_pressure * 2.0 * geom.throat_height())
        elif name == "exit_pressure":
            scale = 1.0 / self.flow.total_pressure
        f64 = dict(dtype=torch.float64, device=self.device)
        rd = self.ops.ref
        J = torch.empty(1, **f64)
        dU = torch.empty_like(u)
        dG = torch.empty((self.ops.n_edges_total, rd.n_qface, 3), **f64)
        work = torch.empty(max(int(self.lib.nz_observables_work_doubles(self._pp)), 1), **f64)
        rc = self.lib.nz_objective(self._pp, self._p(u), oid, scale, self._p(work), self._p(J),
                                   self._p(dU), self._p(dG), self._stream())
        if rc != 0:
            raise _lib.HipError(f"nz_objective failed ({rc}): "
                                f"{self.lib.nz_observables_last_error().decode(errors='replace')}")
        return float(J.item()), dU, dG

    def limit(self, U):
        """Limited copy of ``U``; activity counters as ``numba_backend.py:265-280``."""
        u, npy = self._in(U, "state", "U")
        out = u.clone()
        if self.limiting:
            self._counters.zero_()
            self._apply_limiter(out, self._stream())
            c = self._counters.cpu().tolist()
            self.n_limited += max(int(c[0]), int(c[2]))
            self.n_mean_repaired += int(c[1])
            self.n_limit_calls += 1
        return self._out(out, npy)

    def _apply_limiter(self, u, st, *, fused_positivity: bool = False) -> None:
        """Superbee (when configured) then positivity, in place on device ``u``
        (``limiter.py:408-429``).  ``fused_positivity``: the stage kernel has
        already applied the positivity limiter."""
        if self.superbee:
            cnt = self._counters.data_ptr() + 2 * 8
            self._chk(self.lib.nz_superbee(self._pp, self._p(u), self.tvb, self.area_total,
                                           self._p(self._sb_work), C.c_void_p(cnt), st),
                      "nz_superbee")
        if self.superbee or not fused_positivity:
            self._chk(self.lib.nz_limit(self._pp, self._p(u), None, self._p(self._counters), st),
                      "nz_limit")

    def limit_flags(self, U):
        """``(U_limited, flags, (n_scaled, n_repaired))`` for parity checks.

        ``flags`` bit 0: the cell mean was repaired; bit 1: a probe violated
        the density or pressure bound.  Counters are not accumulated here.
        """
        u, npy = self._in(U, "state", "U")
        out = u.clone()
        cnt = self._torch.zeros(2, dtype=self._torch.int64, device=self.device)
        flags = self._torch.zeros(self.E, dtype=self._torch.uint8, device=self.device)
        if self.limiting:
            self._chk(self.lib.nz_limit(self._pp, self._p(out), self._p(flags), self._p(cnt),
                                        self._stream()), "nz_limit")
        c = cnt.cpu().tolist()
        return self._out(out, npy), self._out(flags, npy), (int(c[0]), int(c[1]))

    def time_step(self, wave_sum):
        if tuple(np.shape(wave_sum)) == (self.E, 1, 1):
            wave_sum = wave_sum.reshape(self.E)
        w, npy = self._in(wave_sum, "elem", "wave_sum")
        dt = self._torch.empty_like(w)
        self._chk(self.lib.nz_time_step(self._pp, self._p(w), self.dt_factor(), self._p(dt),
                                        self._stream()), "nz_time_step")
        dt = dt.reshape(-1, 1, 1)
        return self._out(dt, npy)

    def norm(self, A) -> float:
        a, _ = self._in(A, "state", "A")
        self._chk(self.lib.nz_norm(self._pp, self._p(a), self._p(self._scratch),
                                   self._p(self._res), self._stream()), "nz_norm")
        return float(self._res.item())

    # ------------------------------------------------------------ integrators
    def euler_step(self, U):
        out, res = self.run(U, 1, "euler")
        return out, res

    def ssprk3_step(self, U):
        out, res = self.run(U, 1, "ssprk3")
        return out, res

    def make_step(self, scheme: str):
        if scheme == "euler":
            return self.euler_step
        if scheme == "ssprk3":
            return self.ssprk3_step
        raise ValueError(f"unknown scheme {scheme!r}; choose from {SCHEMES}")

    def _stage_chain(self, n_stages: int) -> list:
        """``nz_problem.bflux_chain`` for a run of ``n_stages`` consecutive
        stages: each stage after the first finds its input's boundary fluxes
        left by the one before (bit 0, no boundary pre-pass), each stage but
        the last leaves its output's for the next (bit 1).  Off with the
        superbee passes, which change a stage's output after it."""
        if not self.chain_bflux or self.superbee:
            return [0] * n_stages
        return [(1 if i > 0 else 0) | (2 if i < n_stages - 1 else 0) for i in range(n_stages)]

    def _stage(self, uin, un, uout, first: bool, a: float, b: float, fac: float, st,
               chain: int = 0):
        self.P.bflux_chain = chain
        try:
            self._chk(self.lib.nz_stage(
                self._pp, self._p(uin), self._p(un) if un is not None else None, self._p(uout),
                self._p(self._dt), fac, a, b, int(first), self._p(self._scratch), None,
                self._p(self._counters), st), "nz_stage")
        finally:
            self.P.bflux_chain = 0
        if self.superbee:
            self._apply_limiter(uout, st, fused_positivity=True)

    def advance(self, cur, nxt, n_steps: int, scheme: str):
        """``n_steps`` steps from device tensor ``cur`` into ``nxt`` (device buffers).

        Returns the buffer that holds the result.  Launches only; no sync.
        """
        if scheme not in SCHEMES:
            raise ValueError(f"unknown scheme {scheme!r}; choose from {SCHEMES}")
        st = self._stream()
        fac = self.dt_factor(scheme)
        t1, t2 = self._S[2], self._S[3]
        per = 1 if scheme == "euler" else 3
        ch = self._stage_chain(n_steps * per)
        for k in range(0, n_steps * per, per):
            if scheme == "euler":
                self._stage(cur, None, nxt, True, 1.0, 1.0, fac, st, ch[k])
            else:
                self._stage(cur, None, t1, True, 1.0, 1.0, fac, st, ch[k])
                self._stage(t1, cur, t2, False, 0.75, 0.25, fac, st, ch[k + 1])
                self._stage(t2, cur, nxt, False, 1.0 / 3.0, 2.0 / 3.0, fac, st, ch[k + 2])
            cur, nxt = nxt, cur
        self._chk(self.lib.nz_norm_finish(self._pp, self._p(self._scratch), self._p(self._res), st),
                  "nz_norm_finish")
        return cur

    def _advance_chunk(self, n_steps: int, scheme: str):
        """``n_steps`` steps from ``_S[0]``; returns the buffer holding the result.

        A chunk length seen before (same scheme, same problem descriptor and
        buffers) is replayed from a HIP graph captured on its second call, so
        a march's ``check_interval`` chunks cost one graph launch each instead
        of one host call per kernel.  The first call of each chunk length runs
        eagerly.  The graph records kernels only (no memset nodes).
        """
        cur, nxt = self._S[0], self._S[1]
        if not self.use_graphs or n_steps < 2:
            return self.advance(cur, nxt, n_steps, scheme)
        key = (n_steps, scheme, bytes(self.P), self._scratch.data_ptr(),
               self.dt_factor(scheme))
        g = self._graphs.get(key)
        if g is None:
            if key not in self._graph_seen:
                self._graph_seen.add(key)
                return self.advance(cur, nxt, n_steps, scheme)
            torch = self._torch
            g = torch.cuda.CUDAGraph()
            torch.cuda.current_stream(self.device).synchronize()
            with torch.cuda.graph(g):
                self.advance(cur, nxt, n_steps, scheme)
            self._graphs[key] = g
        g.replay()
        return cur if n_steps % 2 == 0 else nxt

    def run(self, U, n_steps: int, scheme: str):
        """Advance ``n_steps`` steps; returns ``(U, residual_norm_of_last_step)``."""
        if scheme not in SCHEMES:
            raise ValueError(f"unknown scheme {scheme!r}; choose from {SCHEMES}")
        u, npy = self._in(U, "state", "U")
        if n_steps <= 0:
            return self._out(u.clone(), npy), float("nan")
        self._S[0].copy_(u)
        self._counters.zero_()
        res_buf = self._advance_chunk(n_steps, scheme)
        out = res_buf.clone()
        res = float(self._res.item())
        if self.limiting:
            # per-call counts are summed on the device over the whole chunk
            c = self._counters.cpu().tolist()
            self.n_limited += max(int(c[0]), int(c[2]))
            self.n_mean_repaired += int(c[1])
            self.n_limit_calls += n_steps * (1 if scheme == "euler" else 3)
        return self._out(out, npy), res

    # ------------------------------------------------------------ adjoint
    def _set_boundary_jacobian(self, pu):
        """Forms the boundary-flux Jacobians of the state at device pointer
        ``pu`` (``nz_boundary_jacobian``) and points ``nz_problem.bjac`` at
        them; the caller resets ``P.bjac`` to None when its march ends."""
        if self.n_boundary_points == 0:
            return
        if getattr(self, "_bjac", None) is None:
            self._bjac = self._torch.empty(self.n_boundary_points * 20, dtype=self._torch.float64,
                                           device=self.device)
        self._chk(self.lib.nz_boundary_jacobian(self._pp, pu, self._p(self._bjac), self._stream()),
                  "nz_boundary_jacobian")
        self.P.bjac = self._bjac.data_ptr()

    def _adj_buffers(self):
        if getattr(self, "_adj", None) is None:
            torch = self._torch
            f64 = dict(dtype=torch.float64, device=self.device)
            nw = int(self.lib.nz_adjoint_work_doubles(self._pp))
            shape = (self.E, self.N, 4)
            self._adj = {
                "work": torch.empty(max(nw, 1), **f64),
                "lam": [torch.empty(shape, **f64) for _ in range(3)],
                "acc": torch.empty(shape, **f64),
                "dt": torch.empty(self.E, **f64),
                "res": torch.empty(1, **f64),
            }
        return self._adj

    def residual_vjp(self, U, lam, *, geometry: bool = False):
        """``(dR/dU)^T lam`` of the unfolded residual at ``U``.

        With ``geometry=True`` also returns the cotangents of the per-point
        metrics ``(E, nqv, 5)``, of the edge normals and Jacobians
        ``(F, nqf, 3)`` and of the back pressure (a float).
        """
        torch = self._torch
        self._check_shape(U, "state", "U")
        self._check_shape(lam, "state", "lam")
        u, npy = self._in(U)
        l, _ = self._in(lam)
        a = self._adj_buffers()
        dU = torch.empty_like(u)
        gm = ge = gp = None
        if geometry:
            rd = self.ops.ref
            gm = torch.zeros((self.E, rd.n_qvol, 5), dtype=torch.float64, device=self.device)
            ge = torch.zeros((self.ops.n_edges_total, rd.n_qface, 3), dtype=torch.float64,
                             device=self.device)
            gp = torch.zeros(max(self.ops.n_boundary * rd.n_qface, 1), dtype=torch.float64,
                             device=self.device)
        P = lambda t: self._p(t) if t is not None else None  # noqa: E731
        self._chk(self.lib.nz_residual_vjp(self._pp, self._p(u), self._p(l), self._p(dU), P(gm),
                                           P(ge), P(gp), self._p(a["work"]), self._stream()),
                  "nz_residual_vjp")
        if not geometry:
            return self._out(dU, npy)
        return (self._out(dU, npy), self._out(gm, npy), self._out(ge, npy), float(gp.sum().item()))

    def tune_vjp(self, U, lam) -> int:
        """Autotuner for the residual-VJP tile shape: times the state-cotangent
        VJP on this mesh with one-wave tiles of 16 elements and with four-wave
        tiles of 64 (128 when N <= 4), keeps the faster for this (element kind,
        geometry order, order, flux) for the rest of the process (so every
        backend of a family, batched or alone, uses the same one), and returns
        it (0 or 1).  The two agree to round-off."""
        if self._vjp_key in _VJP_KERNEL:
            self.P.vjp_kernel = _VJP_KERNEL[self._vjp_key]
            return self.P.vjp_kernel
        torch = self._torch
        self._check_shape(U, "state", "U")
        self._check_shape(lam, "state", "lam")
        u, _ = self._in(U)
        lm, _ = self._in(lam)
        a = self._adj_buffers()
        out = torch.empty_like(u)
        st = torch.cuda.current_stream(self.device)
        cost = []
        for choice in (0, 1):
            self.P.vjp_kernel = choice

            def vjp():
                self._chk(self.lib.nz_residual_vjp(self._pp, self._p(u), self._p(lm), self._p(out),
                                                   None, None, None, self._p(a["work"]),
                                                   self._stream()), "nz_residual_vjp")

            for _ in range(2):
                vjp()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(5):
                vjp()
            e1.record(st)
            e1.synchronize()
            cost.append(e0.elapsed_time(e1))
        best = int(cost[1] < cost[0])
        _VJP_KERNEL[self._vjp_key] = best
        self.P.vjp_kernel = best
        return best

    def _frozen_step(self, u):
        """The adjoint's local pseudo-time step: the forward step of the
        converged state with the ``ssprk3`` calibration and without the
        positivity factor (``src/sensitivity.py:330-343``), into the adjoint
        buffers."""
        torch = self._torch
        a = self._adj_buffers()
        st = self._stream()
        Fr = torch.empty_like(u)
        wave = torch.empty(self.E, dtype=torch.float64, device=self.device)
        self._chk(self.lib.nz_rate(self._pp, self._p(u), self._p(Fr), self._p(wave), st), "nz_rate")
        fac = step_coefficient(self.ops.ref.order, self.opts.cfl, "ssprk3")
        self._chk(self.lib.nz_time_step(self._pp, self._p(wave), fac, self._p(a["dt"]), st),
                  "nz_time_step")
        return a

    def adjoint_solve(self, U, seed, *, tolerance: float = 1e-8, max_iterations: int = 100_000,
                      check_interval: int = 100, verbose: bool = False, method: str = "rk4",
                      restart: int = 60):
        """Solve ``(dR/dU)^T lam = seed``; converged when
        ``rms(vjp(lam) - seed) <= tolerance * rms(seed)``.

        ``method="rk4"`` is the reference's pseudo-time march
        (``src/sensitivity.py:289-383``): classical RK4 with the frozen local
        step of the converged state, sized with the ``ssprk3`` calibration and
        without the positivity factor, tested every ``check_interval`` steps;
        ``iterations`` counts steps (four VJPs each).

        ``method="gmres"`` solves the same linear system by restarted GMRES
        (``restart`` vectors), right-preconditioned with that same local step
        times the inverse mass matrix, so ``D = dt M^-1`` is the operator the
        march integrates; GMRES minimises the true residual, whose norm is
        estimated every 8 VJPs (the Arnoldi process stays on the device in
        between) and recomputed at each restart.
        ``iterations`` counts VJPs.  Same fixed point, same test: the
        gradients agree with the march's to the tolerance.

        Returns ``(lam, relative_residual, iterations, converged)``.
        """
        if max_iterations < 1 or check_interval < 1:
            raise ValueError("max_iterations and check_interval must be at least 1")
        if method not in ("rk4", "gmres"):
            raise ValueError(f"unknown adjoint method {method!r}; choose rk4 or gmres")
        torch = self._torch
        self._check_shape(U, "state", "U")
        self._check_shape(seed, "state", "seed")
        u, npy = self._in(U)
        s, _ = self._in(seed)
        self.tune_vjp(u, s)
        a = self._frozen_step(u)
        st = self._stream()
        if method == "gmres":
            lam, rel, its, ok = self._adjoint_gmres(u, s, a, tolerance, max_iterations, restart,
                                                    verbose)
            return self._out(lam, npy), rel, its, ok
        lam, tmp, nxt = a["lam"]
        lam.zero_()
        norm0 = float(torch.sqrt(torch.mean(s * s)).item())
        if norm0 == 0.0:
            return self._out(lam.clone(), npy), 0.0, 0, True
        work, acc, dt = self._p(a["work"]), self._p(a["acc"]), self._p(a["dt"])
        pu, ps = self._p(u), self._p(s)
        bufs = [lam, nxt]  # the step's input and output swap roles every step

        def march(n):
            stream = self._stream()  # the capture stream while a graph records
            for _ in range(n):
                l_, n_ = bufs
                for stage, (src, dst) in enumerate(((l_, tmp), (tmp, n_), (n_, tmp), (tmp, n_)),
                                                   start=1):
                    self._chk(self.lib.nz_adjoint_rk4_stage(
                        self._pp, pu, self._p(src), self._p(l_), ps, dt, acc, self._p(dst),
                        stage, work, stream), "nz_adjoint_rk4_stage")
                bufs.reverse()

        # U is fixed for the whole march: its boundary-flux Jacobians are
        # formed once and the stage launches apply them (nz_problem.bjac)
        # instead of running a boundary pre-pass each.
        self._set_boundary_jacobian(pu)
        try:
            # An even chunk leaves the buffers in their roles, so from its second
            # occurrence it is replayed from a captured HIP graph (kernels only).
            graphs: dict = {}
            it, res, converged = 0, norm0, False
            while it < max_iterations:
                steps = min(check_interval, max_iterations - it)
                if self.use_graphs and steps >= 2 and steps % 2 == 0:
                    g = graphs.get(steps)
                    if g is None and ("seen", steps) not in graphs:
                        graphs[("seen", steps)] = True
                        march(steps)
                    else:
                        if g is None:
                            torch.cuda.current_stream(self.device).synchronize()
                            g = torch.cuda.CUDAGraph()
                            with torch.cuda.graph(g):
                                march(steps)
                            graphs[steps] = g
                        g.replay()
                else:
                    march(steps)
                lam, nxt = bufs
                it += steps
                self._chk(self.lib.nz_adjoint_defect(self._pp, pu, self._p(lam), ps, work,
                                                     self._p(self._scratch), self._p(a["res"]), st),
                          "nz_adjoint_defect")
                res = float(a["res"].item())
                if verbose:
                    print(f"    adjoint {it:>7d}   {res:.4e}  ({res / norm0:.2e} relative)")
                if not math.isfinite(res):
                    break
                if res <= tolerance * norm0:
                    converged = True
                    break
        finally:
            self.P.bjac = None
        a["lam"] = [lam, tmp, nxt]
        return self._out(lam.clone(), npy), res / norm0, it, converged

    def _adjoint_gmres(self, u, s, a, tolerance, max_iterations, restart, verbose, check=8):
        """Restarted GMRES on ``(dR/dU)^T D y = seed``, ``lam = D y``, with
        ``D = dt M^-1`` (the frozen local step), classical Gram-Schmidt with
        one reorthogonalisation.  The Arnoldi process runs on the device: the
        Hessenberg matrix accumulates in device memory and the host reads it
        once every ``check`` VJPs (and at the end of a cycle) to solve the
        small least-squares problem by Givens rotations, so the VJPs between
        reads are queued without waiting for the host."""
        torch = self._torch
        st = self._stream()
        work, pu = self._p(a["work"]), self._p(u)
        dt = a["dt"].view(-1, 1, 1)
        m = max(int(restart), 1)

        def jt(v):  # (dR/dU)^T v
            out = torch.empty_like(v)
            self._chk(self.lib.nz_residual_vjp(self._pp, pu, self._p(v), self._p(out), None, None,
                                               None, work, st), "nz_residual_vjp")
            return out

        def prec(v):  # dt M^-1 v
            out = torch.empty_like(v)
            self._chk(self.lib.nz_inverse_mass(self._pp, self._p(v), self._p(out), st),
                      "nz_inverse_mass")
            return out.mul_(dt)

        lam = torch.zeros_like(s)
        bnorm = float(torch.linalg.vector_norm(s).item())
        if bnorm == 0.0:
            return lam, 0.0, 0, True
        V = torch.empty((m + 1, s.numel()), dtype=torch.float64, device=s.device)
        H = torch.zeros((m + 1, m), dtype=torch.float64, device=s.device)
        r, its, rel = s.clone(), 0, 1.0
        while True:
            beta = float(torch.linalg.vector_norm(r).item())
            rel = beta / bnorm
            if verbose:
                print(f"    adjoint gmres {its:>7d}   {rel:.2e} relative (true residual)")
            if not math.isfinite(rel) or rel <= tolerance or its >= max_iterations:
                break
            V[0] = r.view(-1) / beta
            H.zero_()
            y = np.zeros(0)
            for j in range(m):
                w = jt(prec(V[j].view_as(s))).view(-1)
                its += 1
                Vj = V[:j + 1]
                h = torch.mv(Vj, w)
                w -= torch.mv(Vj.T, h)
                h2 = torch.mv(Vj, w)
                w -= torch.mv(Vj.T, h2)
                hn = torch.linalg.vector_norm(w)
                H[:j + 1, j] = h + h2
                H[j + 1, j] = hn
                V[j + 1] = w / torch.where(hn > 0.0, hn, torch.ones_like(hn))
                k = j + 1
                if k % check == 0 or k == m or its >= max_iterations:
                    y, res = _least_squares_hessenberg(H[:k + 1, :k].cpu().numpy(), beta)
                    if res <= tolerance * bnorm or len(y) < k or its >= max_iterations:
                        break
            yt = torch.from_numpy(y).to(s.device)
            lam += prec((V[:len(y)].T @ yt).view_as(s))
            r = s - jt(lam)
        return lam, rel, its, bool(math.isfinite(rel) and rel <= tolerance)
