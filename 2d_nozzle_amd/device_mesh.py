"""Mesh connectivity and operators built on the device.

The reference builds both on the host, once per mesh (``build_edges``,
``src/geometry/mesh.py:77-208``; ``build_operators``,
``src/spatial/operators.py:209-342``); in a sweep only the node coordinates
change between designs of one grid (``src/api.py:213-244``).  Here:

* :class:`DeviceTopology` runs ``nz_build_edges`` once per grid: the face
  maps the element kernel reads stay on the device, and the reference's
  ``EdgeConnectivity`` (bit-identical to ``build_edges``) is copied back
  only when a host caller asks for it;
* :class:`DeviceOperators` runs ``nz_build_geometry`` for one set of node
  coordinates on a topology (so designs of one grid reuse it): metrics,
  areas, cell-average weights, inverse mass matrices, edge normals and
  Jacobians, in the element kernel's layout, plus the reference's
  ``Operators`` fields, also copied back on demand.  ``HipBackend`` takes a
  ``DeviceOperators`` wherever it takes host operators.

Device memory comes from torch; the library allocates nothing.
"""

from __future__ import annotations

import ctypes as C
from functools import cached_property

import numpy as np

from . import _lib
from . import reference_element as el
from .mesh import EdgeConnectivity, _cell_offsets
from .reference_element import reference_data


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("device mesh building needs a ROCm GPU")
    return torch


def face_ends(kind: str, geometry_order: int) -> list[int]:
    """Local node indices of the two corners of each face, face by face (the
    element-local order ``build_edges`` keys faces by)."""
    vert = el.vertex_indices(kind, geometry_order)
    out = []
    for f in range(el.n_edges(kind)):
        a, b = el.edge_vertex_local_indices(kind, f)
        out += [int(vert[a]), int(vert[b])]
    return out + [0] * (8 - len(out))


def _mesh_check(rc: int, what: str, lib) -> None:
    if rc != 0:
        raise _lib.HipError(f"{what} failed ({rc}): "
                            f"{lib.nz_mesh_last_error().decode(errors='replace')}")


class DeviceTopology:
    """Topology of one conforming mesh with its edges built on the device.

    Mirrors the reference's ``MeshTopology`` fields (``kind``,
    ``geometry_order``, ``elem_nodes``, ``n_nodes``, ``edges``, ``logical``)
    so host code reads it unchanged; ``edges`` is copied back on first use.
    """

    def __init__(self, kind: str, elem_nodes, geometry_order: int, tag_coords, tag_length: float,
                 tag_tol: float, *, n_nodes: int | None = None, logical=None, device=None):
        torch = _torch()
        self.lib = _lib.load()
        self.kind, self.geometry_order = kind, int(geometry_order)
        self.elem_nodes = np.ascontiguousarray(np.asarray(elem_nodes, dtype=np.int64))
        self.n_nodes = int(n_nodes if n_nodes is not None else self.elem_nodes.max() + 1)
        self.logical = dict(logical or {})
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        E, nf = self.elem_nodes.shape[0], el.n_edges(kind)
        dev = dict(device=self.device)
        i64 = dict(dtype=torch.int64, **dev)
        i32 = dict(dtype=torch.int32, **dev)
        n = E * nf
        self.d_elem_nodes = torch.from_numpy(self.elem_nodes).to(self.device)
        tag = np.ascontiguousarray(np.asarray(tag_coords, dtype=np.float64).reshape(-1, 2))
        d_tag = torch.from_numpy(tag).to(self.device)
        m = _lib.MeshDesc()
        m.n_elem, m.n_nodes, m.nface = E, self.n_nodes, nf
        m.nodes_per_elem = self.elem_nodes.shape[1]
        m.elem_nodes = C.c_void_p(self.d_elem_nodes.data_ptr())
        for k, v in enumerate(face_ends(kind, self.geometry_order)):
            m.face_ends[k] = v
        m.tag_coords = C.c_void_p(d_tag.data_ptr())
        m.tag_length, m.tag_tol = float(tag_length), float(tag_tol)
        a = {
            "iedge_elem": torch.empty(max(n // 2, 1) * 2, **i64),
            "iedge_face": torch.empty(max(n // 2, 1) * 2, **i64),
            "bedge_elem": torch.empty(n, **i64), "bedge_face": torch.empty(n, **i64),
            "bedge_tag": torch.empty(n, **i64),
            "face_edge": torch.empty(n, **i64), "face_side": torch.empty(n, **i64),
            "face_neighbour": torch.empty(n, **i64),
            "edge_elem": torch.empty(n, **i64), "edge_face": torch.empty(n, **i64),
            "fe_nbr": torch.empty(n, **i32), "fe_info": torch.empty(n, **i32),
            "bedge": torch.empty(2 * n, **i32), "counts": torch.zeros(8, **i64),
        }
        o = _lib.EdgeMaps()
        for k, t in a.items():
            setattr(o, k, C.c_void_p(t.data_ptr()))
        nbytes = int(self.lib.nz_build_edges_workspace_bytes(E, nf))
        if nbytes < 0:
            raise ValueError("invalid mesh for nz_build_edges")
        work = torch.empty(nbytes, dtype=torch.uint8, **dev)
        st = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _mesh_check(self.lib.nz_build_edges(C.byref(m), C.byref(o), C.c_void_p(work.data_ptr()),
                                            nbytes, st), "nz_build_edges", self.lib)
        counts = a["counts"].cpu().numpy()
        del work
        if counts[6]:
            raise ValueError("non-conforming mesh: a face is shared by more than two elements")
        self.n_interior, self.n_boundary = int(counts[0]), int(counts[1])
        self.tag_counts = {t: int(counts[1 + t]) for t in (1, 2, 3, 4)}
        F = self.n_interior + self.n_boundary
        self.d = a
        # views sized to the actual edge counts
        self.d_edge_elem = a["edge_elem"][:F]
        self.d_edge_face = a["edge_face"][:F]
        self.d_fe_nbr, self.d_fe_info = a["fe_nbr"], a["fe_info"]
        self.d_bedge = a["bedge"][: 2 * max(self.n_boundary, 1)]

    @property
    def n_elem(self) -> int:
        return int(self.elem_nodes.shape[0])

    @property
    def n_faces(self) -> int:
        return el.n_edges(self.kind)

    def tag_slice(self, tag) -> slice:
        return self.edges.tag_slices.get(int(tag), slice(0, 0))

    @cached_property
    def edges(self) -> EdgeConnectivity:
        """The reference's ``EdgeConnectivity``, copied back from the device."""
        a, E, nf = self.d, self.n_elem, self.n_faces
        ni, nb = self.n_interior, self.n_boundary
        host = {k: a[k].cpu().numpy() for k in ("iedge_elem", "iedge_face", "bedge_elem",
                                                "bedge_face", "bedge_tag", "face_edge",
                                                "face_side", "face_neighbour")}
        tag_slices, pos = {}, 0
        for t in (1, 2, 3, 4):
            c = self.tag_counts[t]
            if c:
                tag_slices[t] = slice(pos, pos + c)
            pos += c
        return EdgeConnectivity(
            iedge_elem=host["iedge_elem"][: 2 * ni].reshape(ni, 2),
            iedge_face=host["iedge_face"][: 2 * ni].reshape(ni, 2),
            bedge_elem=host["bedge_elem"][:nb], bedge_face=host["bedge_face"][:nb],
            bedge_tag=host["bedge_tag"][:nb],
            face_edge=host["face_edge"].reshape(E, nf), face_side=host["face_side"].reshape(E, nf),
            face_neighbour=host["face_neighbour"].reshape(E, nf), tag_slices=tag_slices,
        )


def structured_topology(kind: str = "tri", *, nx: int = 14, nr: int = 5, refine: int = 0,
                        geometry_order: int = 1, device=None) -> tuple[DeviceTopology, np.ndarray,
                                                                       np.ndarray]:
    """The reference's structured nozzle grid (``structured_nozzle_topology``,
    ``src/geometry/mesh.py:361-439``) with its edges built on the device;
    returns ``(topology, x_frac, r_frac)`` like ``mesh.structured_nozzle_topology``."""
    q = geometry_order
    ncx, ncy = nx * 2**refine, nr * 2**refine
    n_ix, n_iy = q * ncx + 1, q * ncy + 1
    ci, cj = np.meshgrid(np.arange(ncx), np.arange(ncy), indexing="ij")
    bi, bj = (ci * q).reshape(-1), (cj * q).reshape(-1)
    elem_nodes = np.concatenate(
        [(bj[:, None] + off[None, :, 1]) * n_ix + (bi[:, None] + off[None, :, 0])
         for off in _cell_offsets(kind, q)], axis=0).astype(np.int64)
    x_frac, r_frac = np.linspace(0.0, 1.0, n_ix), np.linspace(0.0, 1.0, n_iy)
    probe = np.column_stack([np.tile(x_frac, n_iy), np.repeat(r_frac, n_ix)])
    tol = 0.25 / max(n_ix, n_iy)
    topo = DeviceTopology(kind, elem_nodes, q, probe, 1.0, tol, n_nodes=n_ix * n_iy,
                          logical={"nx_cells": ncx, "nr_cells": ncy, "nx_nodes": n_ix,
                                   "nr_nodes": n_iy, "refine": refine, "nx_base": nx,
                                   "nr_base": nr}, device=device)
    return topo, x_frac, r_frac


def _geometry_tables(rd) -> tuple[np.ndarray, dict]:
    parts = {
        "dgeom_dxi_vol": rd.dgeom_dxi_vol, "dgeom_deta_vol": rd.dgeom_deta_vol,
        "geom_vol": rd.geom_vol, "w_vol": rd.w_vol, "phi_vol": rd.phi_vol,
        "dgeom_dsigma_face": rd.dgeom_dsigma_face, "geom_face": rd.geom_face,
        "w_face": rd.w_face,
    }
    off, chunks, pos = {}, [], 0
    for k, v in parts.items():
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        off[k], pos = pos, pos + v.size
        chunks.append(v)
    return np.concatenate(chunks), off


class DeviceOperators:
    """Operators of one mesh and order, computed on the device from node
    coordinates and a :class:`DeviceTopology` (field names of the
    reference's ``Operators``, ``src/spatial/operators.py:138-206``).

    ``device_arrays`` holds the element kernel's layout (``geo``, ``area``,
    ``minv``, ``mean_w``, ``edge_geo``, ``fe_nbr``, ``fe_info``, ``bedge``);
    the reference-layout fields are device tensors under ``device_fields``
    and NumPy arrays, copied back on first access, under their own names.
    """

    def __init__(self, topology: DeviceTopology, node_coords, order: int, *,
                 check_jacobian: bool = True, reference_fields: bool = True):
        torch = _torch()
        lib = _lib.load()
        self.topology = topology
        self.ref = rd = reference_data(topology.kind, order, topology.geometry_order)
        self.device = topology.device
        X = node_coords
        if isinstance(X, torch.Tensor):
            self._X_dev = X.to(device=self.device, dtype=torch.float64).contiguous()
        else:
            self._X_dev = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64)).to(self.device)
        self._X_host = None if isinstance(X, torch.Tensor) else np.asarray(X, dtype=np.float64)
        E, N, nqv, nqf = topology.n_elem, rd.n_basis, rd.n_qvol, rd.n_qface
        F = topology.n_interior + topology.n_boundary
        affine = topology.kind == "tri" and topology.geometry_order == 1
        straight = topology.geometry_order == 1
        f64 = dict(dtype=torch.float64, device=self.device)
        tables, off = _geometry_tables(rd)
        self._tables = torch.from_numpy(tables).to(self.device)
        g = _lib.GeometryDesc()
        g.n_elem, g.n_edge = E, F
        g.nq, g.nqv, g.nqf, g.nbf, g.nface = topology.elem_nodes.shape[1], nqv, nqf, N, topology.n_faces
        g.affine, g.straight = int(affine), int(straight)
        g.elem_nodes = C.c_void_p(topology.d_elem_nodes.data_ptr())
        g.coords = C.c_void_p(self._X_dev.data_ptr())
        g.edge_elem = C.c_void_p(topology.d_edge_elem.data_ptr())
        g.edge_face = C.c_void_p(topology.d_edge_face.data_ptr())
        g.tables = C.c_void_p(self._tables.data_ptr())
        for k, v in off.items():
            setattr(g.off, k, v)
        k_arr = {
            "geo": torch.empty(E * 5 if affine else E * nqv * 5, **f64),
            "area": torch.empty(E, **f64),
            "minv": None if affine else torch.empty(E * N * (N + 1) // 2, **f64),
            "mean_w": None if affine else torch.empty(E * N, **f64),
            "edge_geo": torch.empty(max(F, 1) * (3 if straight else 3 * nqf), **f64),
        }
        r_arr = {}
        if reference_fields:
            r_arr = {
                "det": torch.empty(E * nqv, **f64), "metrics": torch.empty(E * nqv * 4, **f64),
                "wdet": torch.empty(E * nqv, **f64), "xy_vol": torch.empty(E * nqv * 2, **f64),
                "mean_w_full": torch.empty(E * N, **f64),
                "normal": torch.empty(max(F, 1) * nqf * 2, **f64),
                "jac": torch.empty(max(F, 1) * nqf, **f64), "length": torch.empty(max(F, 1), **f64),
                "xy_face": torch.empty(max(F, 1) * nqf * 2, **f64),
            }
        bad = torch.zeros(1, dtype=torch.int64, device=self.device)
        o = _lib.GeometryOut()
        for k, t in {**k_arr, **r_arr}.items():
            setattr(o, k, C.c_void_p(t.data_ptr() if t is not None else 0))
        o.bad_jacobian = C.c_void_p(bad.data_ptr())
        st = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _mesh_check(lib.nz_build_geometry(C.byref(g), C.byref(o), st), "nz_build_geometry", lib)
        nbad = int(bad.item())
        if check_jacobian and nbad:
            raise ValueError(f"{nbad} quadrature point(s) have a non-positive Jacobian determinant; "
                             "the mesh is inverted or tangled")
        self.device_arrays = {
            **k_arr, "fe_nbr": topology.d_fe_nbr, "fe_info": topology.d_fe_info,
            "bedge": topology.d_bedge,
        }
        self.device_fields = r_arr
        self.affine, self.straight = affine, straight

    # ---------------------------------------------------------- sizes
    @property
    def n_elem(self) -> int:
        return self.topology.n_elem

    @property
    def n_interior(self) -> int:
        return self.topology.n_interior

    @property
    def n_boundary(self) -> int:
        return self.topology.n_boundary

    @property
    def n_edges_total(self) -> int:
        return self.n_interior + self.n_boundary

    @property
    def elem_area_total(self) -> float:
        return float(self.device_arrays["area"].sum().item())

    # ------------------------------------------------- host views, on demand
    def _host(self, name, shape):
        t = self.device_fields.get(name)
        if t is None:
            raise AttributeError(f"{name} was not computed (reference_fields=False)")
        return t.cpu().numpy().reshape(shape)

    @cached_property
    def node_coords(self) -> np.ndarray:
        return self._X_host if self._X_host is not None else self._X_dev.cpu().numpy()

    @cached_property
    def det_jac(self):
        return self._host("det", (self.n_elem, self.ref.n_qvol))

    @cached_property
    def weighted_det(self):
        return self._host("wdet", (self.n_elem, self.ref.n_qvol))

    @cached_property
    def inverse_jacobian(self):
        """(E, nqv, 4): d xi/dx, d xi/dy, d eta/dx, d eta/dy."""
        return self._host("metrics", (self.n_elem, self.ref.n_qvol, 4))

    @cached_property
    def elem_area(self):
        return self.device_arrays["area"].cpu().numpy()

    @cached_property
    def mean_weights(self):
        return self._host("mean_w_full", (self.n_elem, self.ref.n_basis))

    @cached_property
    def xy_vol(self):
        return self._host("xy_vol", (self.n_elem, self.ref.n_qvol, 2))

    @cached_property
    def edge_normal(self):
        return self._host("normal", (self.n_edges_total, self.ref.n_qface, 2))

    @cached_property
    def edge_jac(self):
        return self._host("jac", (self.n_edges_total, self.ref.n_qface))

    @cached_property
    def edge_length(self):
        return self._host("length", (self.n_edges_total,))

    @cached_property
    def xy_face(self):
        return self._host("xy_face", (self.n_edges_total, self.ref.n_qface, 2))

    @cached_property
    def face_edge(self):
        return self.topology.edges.face_edge

    @cached_property
    def face_sign(self):
        return np.where(self.topology.edges.face_side == 0, 1.0, -1.0)

    @cached_property
    def inv_mass_packed(self):
        """(E, N (N + 1) / 2) upper triangles of the inverse mass matrices
        (non-affine elements), row-major."""
        m = self.device_arrays["minv"]
        if m is None:
            return None
        N = self.ref.n_basis
        return m.cpu().numpy().reshape(self.n_elem, N * (N + 1) // 2)

    # the dense reference arrays are never built on the device
    grad_x = grad_y = grad_x_m = grad_y_m = face_basis = face_basis_m = inv_mass = None


def build_nozzle_operators(geom, *, kind: str = "tri", nx: int = 14, nr: int = 5,
                           refine: int = 0, geometry_order: int = 1, order: int = 1,
                           topology: DeviceTopology | None = None, x_spacing: str = "throat",
                           cluster_strength: float = 3.0, cluster_width: float = 0.12,
                           device=None):
    """``(ops, topology, grids)`` of a nozzle design (the grid of
    ``mesh.build_nozzle_mesh``) with edges and operators built on the device.
    Pass the ``topology`` of an earlier design on the same grid to reuse it:
    then only the node coordinates (host, O(nodes)) and the geometry (device)
    are recomputed."""
    from .mesh import nozzle_node_coords, nozzle_x_distribution

    if topology is None:
        topology, _, r_frac = structured_topology(kind, nx=nx, nr=nr, refine=refine,
                                                  geometry_order=geometry_order, device=device)
    else:
        lg = topology.logical
        if (topology.kind, topology.geometry_order, lg["nx_base"], lg["nr_base"],
                lg["refine"]) != (kind, geometry_order, nx, nr, refine):
            raise ValueError("the topology passed in is of another grid")
        r_frac = np.linspace(0.0, 1.0, lg["nr_nodes"])
    x_frac = nozzle_x_distribution(
        topology.logical["nx_nodes"], nx_base=nx, n_sub=geometry_order * 2**refine,
        spacing=x_spacing, x_throat=geom.throat_location() / geom.length,
        cluster_strength=cluster_strength, cluster_width=cluster_width)
    coords = nozzle_node_coords(x_frac, r_frac, geom.params(), geom.contour)
    return (DeviceOperators(topology, coords, order), topology,
            {"x_frac": x_frac, "r_frac": r_frac})


def build_channel_operators(ncx: int, ncy: int, order: int, *, kind: str = "tri",
                            geometry_order: int = 1, device=None):
    """The synthetic channel of ``mesh.synthetic_channel`` with edges and
    operators built on the device."""
    topology, x, r = structured_topology(kind, nx=ncx, nr=ncy, refine=0,
                                         geometry_order=geometry_order, device=device)
    yw = 1.0 + 1.5 * (x - 0.3) ** 2
    X = np.broadcast_to(x[None, :], (r.size, x.size)).reshape(-1)
    Y = (r[:, None] * yw[None, :]).reshape(-1)
    return DeviceOperators(topology, np.stack([X, Y], axis=-1), order)


class _FamilyTopology:
    """The host-side topology fields a family backend reads."""

    def __init__(self, kind, geometry_order, n_faces, elem_nodes):
        self.kind, self.geometry_order, self.n_faces = kind, geometry_order, n_faces
        self.elem_nodes = elem_nodes


class DeviceFamily:
    """Several designs' ``DeviceOperators`` of one (element kind, order,
    geometry order) as one block-diagonal problem on the device, in the
    layout ``batch.concat_operators`` gives host operators: elements and
    nodes design after design, then the interior edges of every design, then
    the boundary edges of every design.  Only device arrays are concatenated
    (offsets applied on the device); nothing is copied back."""

    def __init__(self, ops_list):
        torch = _torch()
        o0 = ops_list[0]
        key = (o0.topology.kind, o0.topology.geometry_order, o0.ref.order)
        for o in ops_list:
            if (o.topology.kind, o.topology.geometry_order, o.ref.order) != key:
                raise ValueError("a family needs one element kind, order and geometry order")
        self.ref = o0.ref
        dev = o0.device
        ne = np.array([o.n_elem for o in ops_list])
        ni = np.array([o.n_interior for o in ops_list])
        nb = np.array([o.n_boundary for o in ops_list])
        nn = np.array([o.topology.n_nodes for o in ops_list])
        self.elem_offset = eoff = np.concatenate([[0], np.cumsum(ne)])
        self.interior_offset = ioff = np.concatenate([[0], np.cumsum(ni)])
        self.boundary_offset = boff = np.concatenate([[0], np.cumsum(nb)])
        noff = np.concatenate([[0], np.cumsum(nn)])
        NI = int(ioff[-1])
        self.n_elem, self.n_interior, self.n_boundary = int(eoff[-1]), NI, int(boff[-1])
        nf = o0.topology.n_faces
        straight = o0.straight
        per_edge = 3 if straight else 3 * o0.ref.n_qface
        fe_nbr, fe_info, bedge, eg_int, eg_bnd = [], [], [], [], []
        for d, o in enumerate(ops_list):
            a = o.device_arrays
            E, nid, nbd = int(ne[d]), int(ni[d]), int(nb[d])
            nbr = a["fe_nbr"][: E * nf].to(torch.int64)
            fe_nbr.append(torch.where(nbr >= 0, nbr + int(eoff[d]), nbr))
            info = a["fe_info"][: E * nf].to(torch.int64)
            edge = info >> 3
            edge = torch.where(edge < nid, edge + int(ioff[d]), edge - nid + NI + int(boff[d]))
            fe_info.append((edge << 3) | (info & 7))
            b = a["bedge"][: 2 * nbd].to(torch.int64).view(-1, 2)
            bedge.append(torch.stack([b[:, 0] + int(eoff[d]), b[:, 1]], dim=1).reshape(-1))
            eg = a["edge_geo"][: (nid + nbd) * per_edge]
            eg_int.append(eg[: nid * per_edge])
            eg_bnd.append(eg[nid * per_edge:])
        cat = lambda xs: torch.cat(xs) if xs else None  # noqa: E731
        self.device_arrays = {
            "geo": cat([o.device_arrays["geo"] for o in ops_list]),
            "area": cat([o.device_arrays["area"] for o in ops_list]),
            "minv": None if o0.device_arrays["minv"] is None
            else cat([o.device_arrays["minv"] for o in ops_list]),
            "mean_w": None if o0.device_arrays["mean_w"] is None
            else cat([o.device_arrays["mean_w"] for o in ops_list]),
            "edge_geo": cat(eg_int + eg_bnd),
            "fe_nbr": cat(fe_nbr).to(torch.int32),
            "fe_info": cat(fe_info).to(torch.int32),
            "bedge": (cat(bedge) if self.n_boundary else torch.zeros(2, dtype=torch.int64,
                                                                      device=dev)).to(torch.int32),
        }
        self.device_fields = {}
        elem_nodes = np.concatenate([o.topology.elem_nodes + noff[d] for d, o in enumerate(ops_list)])
        self.topology = _FamilyTopology(*key[:2], nf, elem_nodes)
        self.node_coords = np.concatenate([np.asarray(o.node_coords) for o in ops_list])
        self.straight = straight
        self.device = dev

    @property
    def n_edges_total(self) -> int:
        return self.n_interior + self.n_boundary

    @property
    def elem_area_total(self) -> float:
        return float(self.device_arrays["area"].sum().item())
