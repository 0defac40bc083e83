"""Host-side logic of ``HipBackend`` that runs without a GPU: the input shape
guard in front of the C ABI (which takes no sizes) and the locality tile
table of the element kernel."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import nz_module


def _bare_backend(E=37, N=6):
    """A HipBackend shell with only the sizes set: enough for the shape guard,
    which must raise before any device call."""
    be = nz_module("backend")
    bk = be.HipBackend.__new__(be.HipBackend)
    bk.E, bk.N = E, N
    bk._torch = None  # any device call would fail with AttributeError, not ValueError
    return bk


@pytest.mark.parametrize("method", ["residual", "rate", "inverse_mass", "limit", "limit_flags",
                                    "norm", "performance"])
def test_state_shape_rejected_before_launch(method):
    bk = _bare_backend()
    for bad in (np.zeros((36, 6, 4)), np.zeros((37, 3, 4)), np.zeros((37, 6)), np.zeros(37 * 24)):
        with pytest.raises(ValueError, match="must have shape"):
            getattr(bk, method)(bad)


def test_run_and_vjp_shapes_rejected_before_launch():
    bk = _bare_backend()
    good, bad = np.zeros((37, 6, 4)), np.zeros((37, 6, 3))
    with pytest.raises(ValueError, match="must have shape"):
        bk.run(bad, 1, "ssprk3")
    with pytest.raises(ValueError, match="U must have shape"):
        bk.residual_vjp(bad, good)
    with pytest.raises(ValueError, match="lam must have shape"):
        bk.residual_vjp(good, bad)
    with pytest.raises(ValueError, match="seed must have shape"):
        bk.adjoint_solve(good, bad)


def test_time_step_shape_rejected_before_launch():
    bk = _bare_backend()
    for bad in (np.zeros(36), np.zeros((37, 2)), np.zeros((37, 6, 4))):
        with pytest.raises(ValueError, match="wave_sum must have shape"):
            bk.time_step(bad)


def test_project_shape_rejected_before_launch():
    bk = _bare_backend()

    class _Ref:
        n_qvol, n_basis = 9, 6

    class _Ops:
        ref = _Ref()

    bk.ops = _Ops()
    with pytest.raises(ValueError, match="values must have shape"):
        bk.project(np.zeros((37, 6, 4)))


@pytest.mark.parametrize("ncx,ncy,domains", [(128, 64, 8), (100, 41, 8), (37, 5, 3), (3, 1, 8)])
def test_locality_tiles_cover_every_tile_once(ncx, ncy, domains):
    """Every 16-element block appears in exactly one workgroup, with its own
    block index as the norm slot, for tile counts that are and are not
    multiples of the number of L2 domains."""
    be, mesh = nz_module("backend"), nz_module("mesh")
    topo, coords = mesh.synthetic_channel(ncx, ncy)
    E = topo.n_elem
    tab = be.locality_tiles(np.asarray(coords), np.asarray(topo.elem_nodes), E, domains)
    T = (E + be.TILE - 1) // be.TILE
    assert tab.shape == (T, 3) and tab.dtype == np.int64
    assert np.array_equal(np.sort(tab[:, 2]), np.arange(T))
    assert np.array_equal(tab[:, 0], tab[:, 2] * be.TILE)
    assert np.array_equal(tab[:, 1], np.minimum(tab[:, 0] + be.TILE, E))
    covered = np.zeros(E, dtype=np.int64)
    for a, b, _ in tab:
        covered[a:b] += 1
    assert (covered == 1).all()


def test_locality_tiles_deal_contiguous_shares_per_domain():
    """Workgroup b runs on L2 domain b % n_domains; each domain's workgroups
    walk one contiguous run of the Morton order."""
    be, mesh = nz_module("backend"), nz_module("mesh")
    topo, coords = mesh.synthetic_channel(64, 32)
    E, nd = topo.n_elem, 8
    tab = be.locality_tiles(np.asarray(coords), np.asarray(topo.elem_nodes), E, nd)
    T = tab.shape[0]
    share = -(-T // nd)
    # recover the Morton rank of each tile from the table of a single domain
    one = be.locality_tiles(np.asarray(coords), np.asarray(topo.elem_nodes), E, 1)
    rank = np.empty(T, dtype=np.int64)
    rank[one[:, 2]] = np.arange(T)
    for d in range(nd):
        mine = rank[tab[d::nd, 2]]
        assert np.array_equal(mine, np.arange(d * share, min((d + 1) * share, T)))


def test_gmres_least_squares_matches_lstsq():
    """The Givens solve of the GMRES least-squares problem against
    numpy.linalg.lstsq on random upper Hessenberg matrices, and the early end
    at an exact Krylov space (a zero subdiagonal entry)."""
    be = nz_module("backend")
    rng = np.random.default_rng(3)
    for k in (1, 2, 5, 8, 30):
        H = np.triu(rng.standard_normal((k + 1, k)), -1)
        beta = 2.5
        y, res = be._least_squares_hessenberg(H, beta)
        rhs = np.zeros(k + 1)
        rhs[0] = beta
        y_ref = np.linalg.lstsq(H, rhs, rcond=None)[0]
        np.testing.assert_allclose(y, y_ref, rtol=1e-10, atol=1e-12)
        assert abs(res - np.linalg.norm(rhs - H @ y_ref)) <= 1e-12 * beta
    H = np.triu(rng.standard_normal((6, 5)), -1)
    H[3, 2] = 0.0  # the Krylov space is exact after three columns
    y, res = be._least_squares_hessenberg(H, 1.0)
    assert len(y) == 3
    rhs = np.zeros(4)
    rhs[0] = 1.0
    np.testing.assert_allclose(y, np.linalg.lstsq(H[:4, :3], rhs, rcond=None)[0], rtol=1e-10)
