"""Patch tables of the element kernel, on the host: every element in exactly
one patch, distinct norm slots, and compact patches on the structured grids
(2 x 4 cell blocks: 12 faces off the patch, 30 flux-owning faces, so the
flux items of a patch fill two wave rounds at four edge points; the ids come
in runs of four consecutive elements)."""

from __future__ import annotations

import dataclasses

import numpy as np

from conftest import nz_module


def _patch_stats(topo, tab):
    be = nz_module("backend")
    E = topo.n_elem
    tab = tab.reshape(-1, be.PATCH_INTS)
    pid = np.full(E, -1, dtype=np.int64)
    for k in range(16):
        ok = tab[:, k] >= 0
        pid[tab[ok, k]] = np.flatnonzero(ok)
    nb = np.asarray(topo.edges.face_neighbour).astype(np.int64)
    bnd = nb == np.arange(E)[:, None]
    off = (~bnd) & (pid[nb] != pid[:, None])
    n = tab.shape[0]
    n_off = np.bincount(pid, weights=off.sum(axis=1), minlength=n)
    n_bnd = np.bincount(pid, weights=bnd.sum(axis=1), minlength=n)
    return pid, n_off, n_off + n_bnd + (3 * tab[:, 16] - n_off - n_bnd) / 2


def test_patch_table_covers_every_element_once():
    be, mesh = nz_module("backend"), nz_module("mesh")
    topo, coords = mesh.synthetic_channel(40, 13)  # E = 1040: a ragged last patch
    tab = be.patch_table(be.patch_order(topo, coords, topo.n_elem))
    assert tab.dtype == np.int32 and tab.shape[1] == be.PATCH_INTS
    ids, cnt = tab[:, :16], tab[:, 16]
    assert np.all((cnt >= 1) & (cnt <= 16)) and cnt.sum() == topo.n_elem
    lane = np.arange(16)[None, :]
    assert np.array_equal(np.sort(ids[lane < cnt[:, None]]), np.arange(topo.n_elem))
    assert np.all(ids[lane >= cnt[:, None]] == -1)
    assert np.array_equal(np.sort(tab[:, 17]), np.arange(tab.shape[0]))
    # dealt across the 8 L2 domains: workgroup b takes the (b // 8)-th patch
    # of domain b % 8's contiguous share
    share = -(-tab.shape[0] // 8)
    b = np.arange(tab.shape[0])
    pos = (b % 8) * share + b // 8
    assert np.array_equal(tab[pos < tab.shape[0], 17], pos[pos < tab.shape[0]])


def test_structured_patches_are_two_by_four_cell_blocks():
    be, mesh = nz_module("backend"), nz_module("mesh")
    topo, coords = mesh.synthetic_channel(64, 32)
    _, n_off, owners = _patch_stats(topo, be.patch_table(be.patch_order(topo, coords, topo.n_elem)))
    print("off-patch faces: mean", n_off.mean(), "max", n_off.max(), "owners max", owners.max())
    assert n_off.max() <= 12 and owners.max() <= 32
    tab = be.patch_table(be.patch_order(topo, coords, topo.n_elem))
    ids = np.sort(tab[:, :16], axis=1)
    runs = (np.diff(ids, axis=1) != 1).sum(axis=1) + 1
    assert runs.max() == 4  # four runs of four consecutive element ids


def test_centroid_patches_without_grid_indices():
    """A mesh without structured-grid indices falls back to the Morton curve
    of the centroids: still a partition of compact patches."""
    be, mesh = nz_module("backend"), nz_module("mesh")
    topo, coords = mesh.synthetic_channel(64, 32)
    topo = dataclasses.replace(topo, logical={})
    tab = be.patch_table(be.patch_order(topo, coords, topo.n_elem))
    pid, n_off, _ = _patch_stats(topo, tab)
    assert np.all(pid >= 0)
    print("off-patch faces: mean", n_off.mean(), "max", n_off.max())
    assert n_off.mean() <= 20
