"""Batched steady solves of many nozzle designs on one GPU (a parameter sweep).

The reference solves a sweep one design at a time, serially with warm starts
or in a process pool (``src/sweep.py:233-265,300-446``).  Here every design is
marched at once:

* Designs that share (element kind, order, geometry order) form a *family*.
  A family's meshes are concatenated into one block-diagonal problem:
  elements and nodes design after design, then the interior edges of every
  design, then the boundary edges of every design (the device path tells
  boundary from interior edges by one threshold).
* One launch per Runge-Kutta stage advances all active designs of a family
  through a tile table whose tiles never straddle two designs, and whose
  tiles for one design are exactly the tiles that design would have alone.
  The per-tile norm partials are therefore identical, and the per-design
  norms (``nz_norm_segments``) are summed in the same order as a lone solve's,
  so a design marched in the batch takes the same iterations and reaches the
  same state, bit for bit, as when it is marched alone.
* The march control is ``march.StopRules`` per design (the reference's rules,
  ``src/timestepping/march.py:125-230``): one step for the initial residual,
  then chunks of ``check_interval`` steps with the non-finite, tolerance,
  divergence and stall checks.  A design
  that stops is copied out at that step and its tiles leave the table, so it
  costs nothing afterwards (per-design convergence masking).
* Each family runs on its own HIP stream, so small families fill the device
  together; the chunk of a family is captured once as a HIP graph and
  replayed while its set of active designs is unchanged.

Host work (meshes, operators, quasi-1D initial states) runs once per design.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np

from .config import FlowConditions, SolverOptions, residual_scale
from .device_mesh import DeviceFamily
from .initial import quasi1d_initial, quasi1d_values_device
from .march import StopRules
from .mesh import NozzleGeometry, build_nozzle_mesh
from .operators import build_operators

TILE = 16  # elements per tile of the element kernel


@dataclass(frozen=True)
class Design:
    """One point of a sweep (``src/sweep.py:233-265``)."""

    area_ratio: float
    kind: str = "tri"
    order: int = 1
    refine: int = 1
    contour: str = "smooth"
    nx: int = 14
    nr: int = 5
    geometry_order: int = 1

    def geometry(self) -> NozzleGeometry:
        return NozzleGeometry(contour=self.contour, area_ratio=self.area_ratio)


def design_sweep(n: int = 256, seed: int = 1) -> list[Design]:
    """The mixed sweep: area ratio U[1.5, 4], refine, order in {1, 2, 3},
    triangles or quadrilaterals, smooth contour (drawn in that order from
    ``default_rng(seed)``)."""
    rng = np.random.default_rng(seed)
    ar = rng.uniform(1.5, 4.0, n)
    refine = rng.integers(1, 4, n)
    order = rng.integers(1, 4, n)
    kind = rng.choice(np.array(["tri", "quad"]), n)
    return [Design(area_ratio=float(ar[i]), kind=str(kind[i]), order=int(order[i]),
                   refine=int(refine[i])) for i in range(n)]


def concat_operators(ops_list):
    """One block-diagonal operator set from several of the same family.

    Returns the attributes the device backend reads, plus the per-design
    element, interior-edge and boundary-edge offsets.
    """
    o0 = ops_list[0]
    rd = o0.ref
    kind, q, nface = o0.topology.kind, o0.topology.geometry_order, o0.topology.n_faces
    for o in ops_list:
        if (o.topology.kind, o.topology.geometry_order, o.ref.order) != (kind, q, rd.order):
            raise ValueError("concat_operators needs one element kind, order and geometry order")
    ne = np.array([o.n_elem for o in ops_list])
    nn = np.array([np.asarray(o.node_coords).shape[0] for o in ops_list])
    ni = np.array([o.n_interior for o in ops_list])
    nb = np.array([o.n_boundary for o in ops_list])
    eoff = np.concatenate([[0], np.cumsum(ne)])
    noff = np.concatenate([[0], np.cumsum(nn)])
    ioff = np.concatenate([[0], np.cumsum(ni)])
    boff = np.concatenate([[0], np.cumsum(nb)])
    NI = int(ioff[-1])

    def edges_of(o, d):
        ed = o.topology.edges
        fe = np.asarray(ed.face_edge)
        nid = int(ni[d])
        fe_g = np.where(fe < nid, ioff[d] + fe, NI + boff[d] + (fe - nid))
        return fe_g

    def split(a, d):
        a = np.asarray(a)
        return a[: ni[d]], a[ni[d]:]

    face_edge = np.concatenate([edges_of(o, d) for d, o in enumerate(ops_list)])
    face_side = np.concatenate([np.asarray(o.topology.edges.face_side) for o in ops_list])
    face_nbr = np.concatenate([np.asarray(o.topology.edges.face_neighbour) + eoff[d]
                               for d, o in enumerate(ops_list)])
    iedge_face = np.concatenate([np.asarray(o.topology.edges.iedge_face) for o in ops_list])
    iedge_elem = np.concatenate([np.asarray(o.topology.edges.iedge_elem) + eoff[d]
                                 for d, o in enumerate(ops_list)])
    bedge_elem = np.concatenate([np.asarray(o.topology.edges.bedge_elem) + eoff[d]
                                 for d, o in enumerate(ops_list)])
    bedge_face = np.concatenate([np.asarray(o.topology.edges.bedge_face) for o in ops_list])
    bedge_tag = np.concatenate([np.asarray(o.topology.edges.bedge_tag) for o in ops_list])
    nrm = [split(o.edge_normal, d) for d, o in enumerate(ops_list)]
    jac = [split(o.edge_jac, d) for d, o in enumerate(ops_list)]
    edge_normal = np.concatenate([a for a, _ in nrm] + [b for _, b in nrm])
    edge_jac = np.concatenate([a for a, _ in jac] + [b for _, b in jac])
    elem_nodes = np.concatenate([np.asarray(o.topology.elem_nodes) + noff[d]
                                 for d, o in enumerate(ops_list)])
    edges = SimpleNamespace(face_edge=face_edge, face_side=face_side, face_neighbour=face_nbr,
                            iedge_face=iedge_face, iedge_elem=iedge_elem, bedge_elem=bedge_elem,
                            bedge_face=bedge_face, bedge_tag=bedge_tag)
    topo = SimpleNamespace(kind=kind, geometry_order=q, n_faces=nface, edges=edges,
                           elem_nodes=elem_nodes)
    minv = [getattr(o, "inv_mass", None) for o in ops_list]
    return SimpleNamespace(
        ref=rd, topology=topo, n_elem=int(eoff[-1]), n_interior=NI, n_boundary=int(boff[-1]),
        n_edges_total=NI + int(boff[-1]),
        node_coords=np.concatenate([np.asarray(o.node_coords) for o in ops_list]),
        elem_area=np.concatenate([np.asarray(o.elem_area) for o in ops_list]),
        mean_weights=np.concatenate([np.asarray(o.mean_weights) for o in ops_list]),
        inv_mass=None if any(m is None for m in minv) else np.concatenate(minv),
        edge_normal=edge_normal, edge_jac=edge_jac,
        elem_offset=eoff, interior_offset=ioff, boundary_offset=boff,
    )


@dataclass
class DesignResult:
    design: Design
    U: np.ndarray
    converged: bool
    iterations: int
    residual: float
    residual_initial: float
    message: str = ""
    operators: object = None
    flow: FlowConditions | None = None
    wall_time: float = float("nan")  # seconds from the batch's start to this design's stop


class _Family:
    """One (kind, order, geometry order) family marched by one backend."""

    def __init__(self, idx, ops_list, U0_list, flow, opts, device, geoms=None):
        import torch

        from .backend import HipBackend

        self.torch = torch
        self.idx = list(idx)
        on_device = getattr(ops_list[0], "device_arrays", None) is not None
        self.cat = DeviceFamily(ops_list) if on_device else concat_operators(ops_list)
        self.bk = HipBackend(self.cat, flow, opts, device=device)
        if self.bk.superbee:
            raise NotImplementedError("the batched march supports limiter 'positivity' or 'none'")
        self.N = self.cat.ref.n_basis
        self.eoff = self.cat.elem_offset
        self.stream = torch.cuda.Stream(device=self.bk.device)
        f64 = dict(dtype=torch.float64, device=self.bk.device)
        self.cur, self.nxt = self.bk._S[0], self.bk._S[1]
        if U0_list is None:
            # quasi-1D point values of every design on the device, projected
            # for the whole family in one launch
            vals = torch.cat([quasi1d_values_device(o, flow, g, self.bk)
                              for o, g in zip(ops_list, geoms)])
            self.cur.copy_(self.bk.project(vals))
        else:
            self.cur.copy_(torch.from_numpy(np.concatenate(U0_list)).to(self.bk.device))
        self.U0 = self.cur.clone()
        self.res_dev = torch.empty(len(self.idx), **f64)
        self.active = list(range(len(self.idx)))  # local design indices
        self.graph = None
        self.graph_n = 0
        self._set_tiles()

    # -------------------------------------------------------------- tiles
    def _set_tiles(self):
        torch = self.torch
        tiles, seg = [], [0]
        for d in self.active:
            a, b = int(self.eoff[d]), int(self.eoff[d + 1])
            for e0 in range(a, b, TILE):
                tiles.append((e0, min(e0 + TILE, b), len(tiles)))
            seg.append(len(tiles))
        dev = self.bk.device
        self.tiles = torch.tensor(np.asarray(tiles, dtype=np.int64).reshape(-1), device=dev)
        self.seg = torch.tensor(np.asarray(seg, dtype=np.int64), device=dev)
        self.counts = torch.tensor([4.0 * (self.eoff[d + 1] - self.eoff[d]) * self.N
                                    for d in self.active], dtype=torch.float64, device=dev)
        if self.bk.patches:  # the family's table decides which designs march
            self.bk.set_patches(False)
        P = self.bk.P
        P.tiles = C.c_void_p(self.tiles.data_ptr()) if tiles else C.c_void_p(0)
        P.n_tiles = len(tiles)
        P.tiles_cover_mesh = 0  # the live designs' tiles only
        if self.bk._scratch.numel() < max(len(tiles), 1024):
            self.bk._scratch = torch.empty(max(len(tiles), 1024), dtype=torch.float64,
                                           device=dev)
        self.graph = None  # the captured chunk used the old table

    # ------------------------------------------------------------ stepping
    def _launch(self, n_steps: int):
        """``n_steps`` steps between the two ping-pong buffers, then the norms."""
        bk, torch = self.bk, self.torch
        scheme = bk.scheme
        fac = bk.dt_factor(scheme)
        st = bk._stream()
        cur, nxt = self.cur, self.nxt
        t1, t2 = bk._S[2], bk._S[3]
        per = 1 if scheme == "euler" else 3
        ch = bk._stage_chain(n_steps * per)
        for k in range(0, n_steps * per, per):
            if scheme == "euler":
                bk._stage(cur, None, nxt, True, 1.0, 1.0, fac, st, ch[k])
            else:
                bk._stage(cur, None, t1, True, 1.0, 1.0, fac, st, ch[k])
                bk._stage(t1, cur, t2, False, 0.75, 0.25, fac, st, ch[k + 1])
                bk._stage(t2, cur, nxt, False, 1.0 / 3.0, 2.0 / 3.0, fac, st, ch[k + 2])
            cur, nxt = nxt, cur
        bk._chk(bk.lib.nz_norm_segments(bk._pp, bk._p(bk._scratch), bk._p(self.seg),
                                        bk._p(self.counts), len(self.active),
                                        bk._p(self.res_dev), st), "nz_norm_segments")
        return n_steps % 2 == 1  # buffers swapped

    def advance(self, n_steps: int, use_graph: bool):
        """Launch ``n_steps`` steps on the family's stream (no sync)."""
        torch = self.torch
        if not self.active:
            return
        # An even chunk leaves the ping-pong buffers in their roles, so its
        # graph can be replayed until the set of active designs changes.
        graph = use_graph and n_steps % 2 == 0
        with torch.cuda.stream(self.stream):
            if not graph:
                swapped = self._launch(n_steps)
            else:
                if self.graph is None or self.graph_n != n_steps:
                    g = torch.cuda.CUDAGraph()
                    self.stream.synchronize()
                    with torch.cuda.graph(g, stream=self.stream):
                        self._launch(n_steps)
                    self.graph, self.graph_n = g, n_steps
                self.graph.replay()
                swapped = False
        if swapped:
            self.cur, self.nxt = self.nxt, self.cur
            self.graph = None  # the captured pointers no longer match the roles

    def norms(self) -> np.ndarray:
        return self.res_dev[: len(self.active)].cpu().numpy()

    def take(self, d: int) -> np.ndarray:
        a, b = int(self.eoff[d]), int(self.eoff[d + 1])
        return self.cur[a:b].cpu().numpy()


class BatchedSteady:
    """Steady solves of a list of designs, marched together on one GPU."""

    def __init__(self, designs, flow: FlowConditions | None = None,
                 opts: SolverOptions | None = None, *, device=None, ops_list=None,
                 U0_list=None, device_setup: bool = False, geoms=None):
        """``device_setup``: build each grid's connectivity once on the device
        and share it among the designs on that grid (the reference rebuilds
        every design's mesh, ``src/api.py:213-244``), each design's operators
        from its node coordinates on the device, and the quasi-1D initial
        states on the device; otherwise meshes, operators and initial states
        are built on the host per design."""
        self.designs = list(designs)
        self.flow = flow or FlowConditions()
        self.opts = opts or SolverOptions()
        t0 = time.perf_counter()
        if ops_list is not None and U0_list is None:
            # device operators given: the initial states are formed on the device
            if geoms is None:
                geoms = [d.geometry() for d in self.designs]
        else:
            geoms = None
        if ops_list is None and device_setup:
            from .device_mesh import build_nozzle_operators

            ops_list, geoms, topos = [], [], {}
            for dsg in self.designs:
                geom = dsg.geometry()
                key = (dsg.kind, dsg.nx, dsg.nr, dsg.refine, dsg.geometry_order)
                ops, topos[key], _ = build_nozzle_operators(
                    geom, kind=dsg.kind, nx=dsg.nx, nr=dsg.nr, refine=dsg.refine,
                    geometry_order=dsg.geometry_order, order=dsg.order, topology=topos.get(key),
                    device=device)
                ops_list.append(ops)
                geoms.append(geom)
            self.n_topologies = len(topos)
        elif ops_list is None:
            ops_list, U0_list = [], []
            for dsg in self.designs:
                geom = dsg.geometry()
                topo, coords, _ = build_nozzle_mesh(geom, kind=dsg.kind, nx=dsg.nx, nr=dsg.nr,
                                                    refine=dsg.refine,
                                                    geometry_order=dsg.geometry_order)
                ops = build_operators(topo, coords, dsg.order, dense=False)
                ops_list.append(ops)
                U0_list.append(quasi1d_initial(ops, self.flow, geom))
        self.ops_list = ops_list
        self.host_setup_s = time.perf_counter() - t0
        fams: dict = {}
        for i, o in enumerate(ops_list):
            key = (o.topology.kind, o.topology.geometry_order, o.ref.order)
            fams.setdefault(key, []).append(i)
        t0 = time.perf_counter()
        self.families = [
            _Family(idx, [ops_list[i] for i in idx],
                    None if U0_list is None else [U0_list[i] for i in idx], self.flow, self.opts,
                    device, geoms=None if geoms is None else [geoms[i] for i in idx])
            for idx in fams.values()]
        if U0_list is None:  # formed on the device: host copies for the results and callers
            U0_list = [None] * len(ops_list)
            for f in self.families:
                u0 = f.U0.cpu().numpy()
                for j, g in enumerate(f.idx):
                    U0_list[g] = u0[int(f.eoff[j]):int(f.eoff[j + 1])]
        self.U0_list = U0_list
        self.device_setup_s = time.perf_counter() - t0

    @property
    def n_elements(self) -> int:
        return int(sum(o.n_elem for o in self.ops_list))

    def solve(self, *, use_graph: bool = True, progress=None,
              concurrent: bool = True) -> list[DesignResult]:
        """March every design to its own stop; results in input order.

        ``concurrent=False`` queues every family on one stream, one after the
        other (the same launches and results; for measuring what running the
        families side by side on their own streams gains)."""
        import torch

        if not concurrent:
            shared = torch.cuda.Stream(device=self.families[0].bk.device)
            for fam in self.families:
                fam.stream = shared
                fam.graph = None
        opts = self.opts
        scale = residual_scale(self.flow, 1.0)
        chunk = min(max(1, opts.check_interval), opts.max_iterations)
        results: list[DesignResult | None] = [None] * len(self.designs)
        state = {}  # global index -> dict(res0, window)

        def finish(fam, d, conv, done, res, msg):
            g = fam.idx[d]
            results[g] = DesignResult(
                design=self.designs[g], U=fam.take(d), converged=conv, iterations=done,
                residual=float(res), residual_initial=float(state[g]["res0"]), message=msg,
                operators=self.ops_list[g], flow=self.flow,
                wall_time=time.perf_counter() - t0)

        def check(first: bool, done: int):
            for fam in self.families:
                if not fam.active:
                    continue
                res = fam.norms()
                keep = []
                for j, d in enumerate(fam.active):
                    g, r = fam.idx[d], float(res[j])
                    if first:
                        state[g] = {"res0": r, "rules": StopRules(opts, scale)}
                        stop = state[g]["rules"].begin(r)
                    else:
                        stop = state[g]["rules"].check(done, r)
                        if stop is None and done >= opts.max_iterations:
                            stop = state[g]["rules"].budget(done)
                    if stop is None:
                        keep.append(d)
                    else:
                        finish(fam, d, stop.converged, done, r, stop.message)
                if keep != fam.active:
                    fam.active = keep
                    fam._set_tiles()

        t0 = time.perf_counter()
        for fam in self.families:
            fam.advance(1, use_graph=False)
        torch.cuda.synchronize()
        done = 1
        check(True, done)
        while any(f.active for f in self.families) and done < opts.max_iterations:
            n = min(chunk, opts.max_iterations - done)
            for fam in self.families:
                fam.advance(n, use_graph)
            torch.cuda.synchronize()
            done += n
            check(False, done)
            if progress is not None:
                progress(done, sum(len(f.active) for f in self.families))
        self.wall_s = time.perf_counter() - t0
        for fam in self.families:  # anything left at max_iterations
            for d in list(fam.active):
                g = fam.idx[d]
                if results[g] is None:
                    finish(fam, d, False, done, float("nan"), "stopped")
        return results


# ------------------------------------------------------------- batched adjoint
@dataclass
class DesignGradient:
    design: Design
    objective: float
    gradient: dict
    adjoint_iterations: int
    adjoint_relative_residual: float
    converged: bool


def _design_grids(dsg: Design):
    """The node fractions of a design's logical grid (the node distributions
    of ``build_nozzle_mesh``, no connectivity) and its parameters."""
    from .mesh import nozzle_x_distribution

    geom = dsg.geometry()
    q, sub = dsg.geometry_order, dsg.geometry_order * 2**dsg.refine
    x_frac = nozzle_x_distribution(q * dsg.nx * 2**dsg.refine + 1, nx_base=dsg.nx, n_sub=sub,
                                   spacing="throat", x_throat=geom.throat_location() / geom.length,
                                   cluster_strength=3.0, cluster_width=0.12)
    r_frac = np.linspace(0.0, 1.0, q * dsg.nr * 2**dsg.refine + 1)
    return x_frac, r_frac, dict(geom.params()), geom.contour


def _design_tangent(dsg: Design, name: str = "area_ratio") -> np.ndarray:
    """Exact ``d node coordinates / d name`` of a design (complex step)."""
    from .mesh import node_coords_tangent

    x_frac, r_frac, p, contour = _design_grids(dsg)
    return node_coords_tangent(x_frac, r_frac, p, contour, name)


def batched_gradients(results: list, flow: FlowConditions, opts: SolverOptions, *,
                      objective: str = "thrust", tolerance: float = 1e-8,
                      max_iterations: int = 100_000, check_interval: int = 100,
                      group_size: int | None = None,
                      device=None) -> list:
    """d objective / d area ratio for every converged design of a batch, from
    steady adjoints marched together (``src/sensitivity.py:289-460``).

    Designs are grouped in families as in the forward batch (optionally split
    into groups of ``group_size`` with similar forward iteration counts), and
    the groups march concurrently, each on its own stream.  Each group's
    adjoint march is the lone one (classical RK4 in pseudo-time with the
    frozen local step of the converged state): element-local updates on a
    block-diagonal mesh, so each design's iterates are those it has alone.
    Every ``check_interval`` steps each design's defect
    ``rms(vjp(lam) - seed) / rms(seed)`` is tested on its own segment; a
    design that passes is copied out at that step and its tiles leave the
    adjoint kernels' tile table, so it costs nothing afterwards.  The
    geometry cotangents of a whole group come from one VJP launch and are
    split back per design for the host chain to node coordinates and the
    design variable.
    """
    import torch

    from .backend import HipBackend
    from .config import step_coefficient
    from .sensitivity import _Throat, edge_vjp, metric_vjp

    def design_objective(g):
        # the device objective and its partials (nz_objective) on the
        # design's own problem, as its lone adjoint (DifferentiableCase) takes
        # them: the same kernel, the same summation order
        r = results[g]
        J, dU, dG = HipBackend(r.operators, flow, opts, device=device).objective(
            r.U, objective, geom=_Throat(r.design.geometry().throat_half_height))
        return J, dU.cpu().numpy(), dG.cpu().numpy()

    out: list = [None] * len(results)
    groups: dict = {}
    for g, r in enumerate(results):
        if r is None or not r.converged:
            continue
        o = r.operators
        groups.setdefault((o.topology.kind, o.topology.geometry_order, o.ref.order), []).append(g)
    # Within a family, designs with similar forward iteration counts (a proxy
    # for the adjoint's) march together in groups, so that a group stops soon
    # after its slowest design instead of the family's.
    parts = []
    for idx in groups.values():
        idx = sorted(idx, key=lambda g: results[g].iterations)
        step = group_size or len(idx)
        parts += [idx[k:k + step] for k in range(0, len(idx), step)]

    def set_vjp_tiles(f):
        # the live designs' elements in tiles of the adjoint kernels' width;
        # a captured chunk used the old table
        te, tiles = f.te, []
        for j, d in enumerate(f.done):
            if not d:
                a, b = int(f.eoff[j]), int(f.eoff[j + 1])
                for e0 in range(a, b, te):
                    tiles += [e0, min(e0 + te, b)]
        f.vt = torch.tensor(tiles if tiles else [0, 0], dtype=torch.int64, device=f.bk.device)
        P = f.bk.P
        P.vjp_tiles = C.c_void_p(f.vt.data_ptr())
        P.n_vjp_tiles = len(tiles) // 2
        f.graph = None

    fams = []
    for idx in parts:
        f = SimpleNamespace(idx=idx)
        ops_g = [results[g].operators for g in idx]
        f.cat = (DeviceFamily(ops_g) if getattr(ops_g[0], "device_arrays", None) is not None
                 else concat_operators(ops_g))
        f.bk = bk = HipBackend(f.cat, flow, opts, device=device)
        if bk.patches:  # patches would mix the designs' elements
            bk.set_patches(False)
        f.stream = torch.cuda.Stream(device=bk.device)
        dev = bk.device
        f.parts = [design_objective(g) for g in idx]
        f.eoff = eoff = f.cat.elem_offset
        f.u = torch.from_numpy(np.concatenate([results[g].U for g in idx])).to(dev)
        f.s = torch.from_numpy(np.concatenate([p[1] for p in f.parts])).to(dev)
        bk.tune_vjp(f.u, f.s)  # the tile shape fixes the tile table's width
        f.te = int(bk.lib.nz_adjoint_tile_elements(bk._pp))
        f.sl = [slice(int(eoff[j]), int(eoff[j + 1])) for j in range(len(idx))]
        # per-design sums over contiguous element segments as a padded gather
        # (graph-capturable, fixed order)
        lens = np.diff(eoff)
        kk = np.arange(int(lens.max()))
        f.seg_idx = torch.tensor(np.where(kk[None, :] < lens[:, None], eoff[:-1, None] + kk[None, :],
                                          0), dtype=torch.int64, device=dev)
        f.seg_mask = torch.tensor(kk[None, :] < lens[:, None], dtype=torch.float64, device=dev)
        f.count = torch.tensor(np.diff(eoff) * f.cat.ref.n_basis * 4.0, dtype=torch.float64,
                               device=dev)
        f.a = bk._adj_buffers()
        f.lam_out = torch.zeros_like(f.u)
        f.dU = torch.empty_like(f.u)
        with torch.cuda.stream(f.stream):
            st = bk._stream()
            Fr = torch.empty_like(f.u)
            wave = torch.empty(f.cat.n_elem, dtype=torch.float64, device=dev)
            bk._chk(bk.lib.nz_rate(bk._pp, bk._p(f.u), bk._p(Fr), bk._p(wave), st), "nz_rate")
            fac = step_coefficient(f.cat.ref.order, opts.cfl, "ssprk3")
            bk._chk(bk.lib.nz_time_step(bk._pp, bk._p(wave), fac, bk._p(f.a["dt"]), st),
                    "nz_time_step")
            f.a["lam"][0].zero_()
            e2 = (f.s * f.s).sum(dim=(1, 2))
            f.norm0 = torch.sqrt((e2[f.seg_idx] * f.seg_mask).sum(dim=1) / f.count).cpu().numpy()
        f.done = [bool(v == 0.0) for v in f.norm0]
        f.its = [0] * len(idx)
        f.rels = [0.0 if d else float("nan") for d in f.done]
        set_vjp_tiles(f)
        with torch.cuda.stream(f.stream):  # the group's fixed states: boundary Jacobians once
            bk._set_boundary_jacobian(bk._p(f.u))
        fams.append(f)

    def launch(f, n):
        # an even chunk leaves the lambda buffers in their roles: capture it
        # once per family and replay it
        if n % 2 == 0:
            if getattr(f, "graph", None) is None or f.graph_n != n:
                f.stream.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=f.stream):
                    launch_eager(f, n)
                f.graph, f.graph_n = g, n
            with torch.cuda.stream(f.stream):
                f.graph.replay()
        else:
            launch_eager(f, n)

    def launch_eager(f, n):
        bk, (lam, tmp, nxt) = f.bk, f.a["lam"]
        work, acc, dtp = bk._p(f.a["work"]), bk._p(f.a["acc"]), bk._p(f.a["dt"])
        pu, ps = bk._p(f.u), bk._p(f.s)
        with torch.cuda.stream(f.stream):
            st = bk._stream()
            for _ in range(n):
                for stage, (li, lo) in enumerate(((lam, tmp), (tmp, nxt), (nxt, tmp), (tmp, nxt)),
                                                 start=1):
                    bk._chk(bk.lib.nz_adjoint_rk4_stage(bk._pp, pu, bk._p(li), bk._p(lam), ps,
                                                        dtp, acc, bk._p(lo), stage, work, st),
                            "rk4")
                lam, nxt = nxt, lam
            f.a["lam"] = [lam, tmp, nxt]
            bk._chk(bk.lib.nz_residual_vjp(bk._pp, pu, bk._p(lam), bk._p(f.dU), None, None, None,
                                           work, st), "nz_residual_vjp")
            d2 = ((f.dU - f.s) ** 2).sum(dim=(1, 2))
            f.res_dev = torch.sqrt((d2[f.seg_idx] * f.seg_mask).sum(dim=1) / f.count)

    it = 0
    while it < max_iterations and any(not all(f.done) for f in fams):
        n = min(check_interval, max_iterations - it)
        live = [f for f in fams if not all(f.done)]
        for f in live:
            launch(f, n)
        it += n
        for f in live:
            f.stream.synchronize()
            res = f.res_dev.cpu().numpy()
            lam = f.a["lam"][0]
            stopped = False
            for j in range(len(f.idx)):
                if f.done[j]:
                    continue
                r = float(res[j])
                if not np.isfinite(r) or r <= tolerance * f.norm0[j]:
                    f.lam_out[f.sl[j]] = lam[f.sl[j]]
                    f.its[j], f.rels[j], f.done[j] = it, r / f.norm0[j], True
                    stopped = True
            if stopped and not all(f.done):
                set_vjp_tiles(f)
    for f in fams:
        lam = f.a["lam"][0]
        for j in range(len(f.idx)):
            if not f.done[j]:
                f.lam_out[f.sl[j]] = lam[f.sl[j]]
                f.its[j] = it
        # geometry cotangents of the whole group, split per design
        bk, cat = f.bk, f.cat
        bk.P.vjp_tiles, bk.P.n_vjp_tiles = C.c_void_p(0), 0
        bk.P.bjac = None
        rd, nqf = cat.ref, cat.ref.n_qface
        dev = bk.device
        gm = torch.zeros((cat.n_elem, rd.n_qvol, 5), dtype=torch.float64, device=dev)
        ge = torch.zeros((cat.n_edges_total, nqf, 3), dtype=torch.float64, device=dev)
        gp = torch.zeros(max(cat.n_boundary * nqf, 1), dtype=torch.float64, device=dev)
        bk._chk(bk.lib.nz_residual_vjp(bk._pp, bk._p(f.u), bk._p(f.lam_out), bk._p(f.dU),
                                       bk._p(gm), bk._p(ge), bk._p(gp), bk._p(f.a["work"]),
                                       bk._stream()), "nz_residual_vjp")
        gm, ge = gm.cpu().numpy(), ge.cpu().numpy()
        ioff, boff, NI = cat.interior_offset, cat.boundary_offset, cat.n_interior
        for j, g in enumerate(f.idx):
            ops = results[g].operators
            ge_d = np.concatenate([ge[ioff[j]:ioff[j + 1]], ge[NI + boff[j]:NI + boff[j + 1]]])
            J, _, J_edge = f.parts[j]
            gX = edge_vjp(ops, J_edge) - metric_vjp(ops, gm[f.sl[j]]) - edge_vjp(ops, ge_d)
            dsg = results[g].design
            dX = _design_tangent(dsg)
            conv = bool(np.isfinite(f.rels[j]) and f.rels[j] <= tolerance)
            out[g] = DesignGradient(design=dsg, objective=J,
                                    gradient={"area_ratio": float((gX * dX).sum())},
                                    adjoint_iterations=f.its[j],
                                    adjoint_relative_residual=f.rels[j], converged=conv)
    return out


# ------------------------------------------------------------- the sweep
def build_device_designs(designs, device=None):
    """Device operators of every design, one connectivity build per grid."""
    from .device_mesh import build_nozzle_operators

    ops_list, topos = [], {}
    for dsg in designs:
        key = (dsg.kind, dsg.nx, dsg.nr, dsg.refine, dsg.geometry_order)
        ops, topos[key], _ = build_nozzle_operators(
            dsg.geometry(), kind=dsg.kind, nx=dsg.nx, nr=dsg.nr, refine=dsg.refine,
            geometry_order=dsg.geometry_order, order=dsg.order, topology=topos.get(key),
            device=device)
        ops_list.append(ops)
    return ops_list


def _serpentine(shape: tuple[int, ...]) -> list[tuple[int, ...]]:
    """Grid indices walked so that consecutive points are neighbours: the
    last axis runs forward and backward alternately."""
    import itertools

    out = list(itertools.product(*[range(n) for n in shape]))
    if len(shape) <= 1:
        return out
    n_last, walked = shape[-1], []
    for start in range(0, len(out), n_last):
        block = out[start:start + n_last]
        walked += block[::-1] if sum(block[0][:-1]) % 2 == 1 else block
    return walked


SWEEP_AXES = ("area_ratio", "order", "refine")


def sweep(*, flow: FlowConditions | None = None, options: SolverOptions | None = None,
          kind: str = "tri", contour: str = "smooth", order: int = 1, refine: int = 1,
          nx: int = 14, nr: int = 5, warm_start: bool = False, device=None, **grids):
    """A parameter sweep on the batched device march, tabulated as the
    reference's ``sweep`` does (``src/sweep.py:300-446``).

    Iterable keywords among ``area_ratio``, ``order`` and ``refine`` are the
    axes of a Cartesian grid; the other settings are fixed.  Every point is
    marched in one batch (grids' connectivity built once on the device,
    operators and initial states on the device), and each converged point's
    metrics come from the device's performance integrals.  ``warm_start``:
    the reference starts each point from the previous one's converged field
    (serpentine order, ``src/sweep.py:379-414``); batched, the even points of
    that order march first from the quasi-1D state and each odd point then
    starts from its converged predecessor on the same grid and order.
    Returns an ``outputs.SweepResult``.
    """
    from .backend import HipBackend
    from .initial import operating_regime, solve_quasi1d
    from .march import probe_minima
    from .outputs import METRICS, SweepResult, sweep_metrics

    flow = flow or FlowConditions()
    options = options or SolverOptions(print_interval=0)
    names = tuple(k for k in grids)
    for k in names:
        if k not in SWEEP_AXES:
            raise NotImplementedError(f"the batched sweep varies {SWEEP_AXES}; {k!r} is not one")
    if not names:
        raise ValueError("sweep needs at least one iterable parameter grid, e.g. "
                         "sweep(area_ratio=np.linspace(2, 4, 5))")
    axes = [np.atleast_1d(np.asarray(list(grids[n]), dtype=float)) for n in names]
    shape = tuple(len(a) for a in axes)
    total = int(np.prod(shape))
    points, designs = [None] * total, [None] * total
    for idx in np.ndindex(*shape):
        fi = int(np.ravel_multi_index(idx, shape))
        pt = {n: float(axes[i][j]) for i, (n, j) in enumerate(zip(names, idx))}
        points[fi] = pt
        designs[fi] = Design(area_ratio=pt.get("area_ratio", 2.5), kind=kind,
                             order=int(pt.get("order", order)), refine=int(pt.get("refine", refine)),
                             contour=contour, nx=nx, nr=nr)
    ops_list = build_device_designs(designs, device)
    results: list = [None] * total
    if not warm_start:
        run = BatchedSteady(designs, flow, options, device=device, ops_list=ops_list)
        results = run.solve()
    else:
        walk = [int(np.ravel_multi_index(i, shape)) for i in _serpentine(shape)]
        first, second = walk[0::2], walk[1::2]
        run1 = BatchedSteady([designs[g] for g in first], flow, options, device=device,
                             ops_list=[ops_list[g] for g in first])
        for g, r in zip(first, run1.solve()):
            results[g] = r
        U0 = []
        for g in second:
            prev = walk[walk.index(g) - 1]
            pr = results[prev]
            same = (designs[prev].kind, designs[prev].order, designs[prev].refine) == (
                designs[g].kind, designs[g].order, designs[g].refine)
            U0.append(pr.U if pr.converged and same else None)
        if second:
            cold = [u is None for u in U0]
            if all(cold):
                run2 = BatchedSteady([designs[g] for g in second], flow, options, device=device,
                                     ops_list=[ops_list[g] for g in second])
            else:
                # initial states: the predecessor's field, or the quasi-1D
                # state (device-projected) where there is none
                from .initial import quasi1d_initial

                for j, g in enumerate(second):
                    if U0[j] is None:
                        bk = HipBackend(ops_list[g], flow, options, device=device)
                        U0[j] = quasi1d_initial(ops_list[g], flow, designs[g].geometry(),
                                                backend=bk)
                run2 = BatchedSteady([designs[g] for g in second], flow, options, device=device,
                                     ops_list=[ops_list[g] for g in second], U0_list=U0)
            for g, r in zip(second, run2.solve()):
                results[g] = r
    values = {m: np.full(total, np.nan) for m in METRICS}
    converged = np.zeros(total, dtype=bool)
    regimes, messages = [""] * total, [""] * total
    for g, r in enumerate(results):
        converged[g], messages[g] = r.converged, r.message
        geom = designs[g].geometry()
        q1d = solve_quasi1d(geom, flow)
        regimes[g] = operating_regime(geom.area_ratio, flow.back_pressure_ratio, flow.gamma)[0].value
        if not r.converged:
            continue
        rmin, pmin = probe_minima(r.U, ops_list[g], flow.gamma)
        r.min_density, r.min_pressure = rmin, pmin
        perf = HipBackend(ops_list[g], flow, options, device=device).performance(
            r.U, geom=geom, quasi1d=q1d)
        for m, v in sweep_metrics(perf, r, geom, flow, q1d).items():
            values[m][g] = v
    return SweepResult(parameters=names, points=points, values=values, converged=converged,
                       regimes=regimes, messages=messages, shape=shape)
