"""Host side of the batched sweep: the block-diagonal concatenation of meshes."""

from __future__ import annotations

import numpy as np

from conftest import nz_module


def _ops(kind, order, refine, ar):
    mesh, op = nz_module("mesh"), nz_module("operators")
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=ar)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=kind, refine=refine)
    return op.build_operators(topo, coords, order, dense=False)


def test_concat_keeps_every_design_intact():
    batch = nz_module("batch")
    parts = [_ops("tri", 2, 0, 2.0), _ops("tri", 2, 1, 3.0), _ops("tri", 2, 0, 1.7)]
    cat = batch.concat_operators(parts)
    assert cat.n_elem == sum(o.n_elem for o in parts)
    assert cat.n_interior == sum(o.n_interior for o in parts)
    ed = cat.topology.edges
    for d, o in enumerate(parts):
        e0, e1 = cat.elem_offset[d], cat.elem_offset[d + 1]
        fe = ed.face_edge[e0:e1]
        ofe = np.asarray(o.topology.edges.face_edge)
        inner = ofe < o.n_interior
        # interior edges first (all designs), then boundary edges (all designs)
        assert np.all(fe[inner] < cat.n_interior)
        assert np.all(fe[~inner] >= cat.n_interior)
        # the remapped edges carry the same geometry
        np.testing.assert_array_equal(cat.edge_normal[fe], np.asarray(o.edge_normal)[ofe])
        np.testing.assert_array_equal(cat.edge_jac[fe], np.asarray(o.edge_jac)[ofe])
        # neighbours stay inside the design
        nb = ed.face_neighbour[e0:e1]
        assert nb.min() >= e0 and nb.max() < e1
        np.testing.assert_array_equal(nb - e0, np.asarray(o.topology.edges.face_neighbour))
        b = slice(cat.n_interior + cat.boundary_offset[d],
                  cat.n_interior + cat.boundary_offset[d + 1])
        np.testing.assert_array_equal(ed.bedge_elem[b.start - cat.n_interior:
                                                    b.stop - cat.n_interior] - e0,
                                      np.asarray(o.topology.edges.bedge_elem))


def test_concat_rejects_mixed_families():
    batch = nz_module("batch")
    import pytest

    with pytest.raises(ValueError, match="one element kind"):
        batch.concat_operators([_ops("tri", 1, 0, 2.0), _ops("tri", 2, 0, 2.0)])


def test_design_sweep_is_seeded_and_mixed():
    batch = nz_module("batch")
    a, b = batch.design_sweep(256, 1), batch.design_sweep(256, 1)
    assert a == b
    assert {d.kind for d in a} == {"tri", "quad"}
    assert {d.order for d in a} == {1, 2, 3} and {d.refine for d in a} == {1, 2, 3}
    ar = np.array([d.area_ratio for d in a])
    assert ar.min() >= 1.5 and ar.max() <= 4.0
