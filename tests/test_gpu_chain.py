"""Boundary fluxes handed from one Runge-Kutta stage to the next
(``nz_problem.bflux_chain``): each stage's epilogue writes the boundary
fluxes of its output for the next stage, which then skips the boundary
pre-pass.

Bar: every march is bit-identical with the chaining on and off (the states,
the residual norm and the limiter counters), on tiles and on patch tiles,
for triangles and quadrilaterals, straight and curved, at P0-P3, with
inflow, outflow (including the backflow blend), wall and axis faces.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import nz_module

pytestmark = pytest.mark.gpu


def _nozzle(kind, order, q, refine=2, flux="roe", contour="smooth", pb=None, limiter="positivity"):
    mesh, op, cf, ini, be = (nz_module(m) for m in ("mesh", "operators", "config", "initial",
                                                   "backend"))
    geom = mesh.NozzleGeometry(contour=contour, area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=kind, refine=refine, geometry_order=q)
    ops = op.build_operators(topo, coords, order, dense=False)
    flow = cf.FlowConditions(flux=flux)
    if pb is not None:
        flow = flow.replace(back_pressure_ratio=pb)
    U0 = ini.quasi1d_initial(ops, flow, geom)
    return be.HipBackend(ops, flow, cf.SolverOptions(limiter=limiter)), U0


def _both(bk, U0, n, scheme):
    """The march with the chaining off, then forced on (whatever the
    backend's size rule picked)."""
    out = []
    for chain in (False, True):
        bk.chain_bflux = chain
        bk.n_limited = 0
        U, res = bk.run(U0, n, scheme)
        out.append((U, res, bk.n_limited))
    return out


@pytest.mark.parametrize("kind,order,q", [("tri", 1, 1), ("tri", 2, 2), ("tri", 3, 1),
                                          ("quad", 2, 2), ("quad", 3, 1), ("tri", 0, 1)])
@pytest.mark.parametrize("scheme", ["ssprk3", "euler"])
def test_chained_march_bit_identical_on_nozzles(kind, order, q, scheme):
    bk, U0 = _nozzle(kind, order, q)
    (Ua, ra, la), (Ub, rb, lb) = _both(bk, U0, 7, scheme)
    print(kind, order, q, scheme, ra, la)
    assert np.array_equal(Ua, Ub) and ra == rb and la == lb


@pytest.mark.parametrize("flux", ["roe", "hllc", "ausmup"])
def test_chained_march_bit_identical_with_backflow(flux):
    """A high back pressure drives the outflow faces into the backflow blend
    while the march runs."""
    bk, U0 = _nozzle("tri", 2, 1, flux=flux, pb=0.95)
    (Ua, ra, _), (Ub, rb, _) = _both(bk, U0, 12, "ssprk3")
    assert np.array_equal(Ua, Ub) and ra == rb


@pytest.mark.parametrize("patches", [False, True])
def test_chained_march_bit_identical_on_the_channel(patches):
    """The P3 channel at 2^14 elements: locality-ordered tiles or patch
    tiles, limiter active."""
    cf, be, wl = nz_module("config"), nz_module("backend"), nz_module("workloads")
    ops, U = wl.channel_case(128, 64, 3)
    bk = be.HipBackend(ops, cf.FlowConditions(), cf.SolverOptions(limiter="positivity"))
    bk.set_patches(patches)
    (Ua, ra, la), (Ub, rb, lb) = _both(bk, U, 5, "ssprk3")
    assert np.array_equal(Ua, Ub) and ra == rb and la == lb


def test_chaining_is_refused_with_superbee():
    """The superbee passes change a stage's output after it, so the backend
    never chains them and the library refuses the flag."""
    bk, U0 = _nozzle("tri", 1, 1, limiter="superbee")
    assert bk._stage_chain(3) == [0, 0, 0]
    bk.P.bflux_chain = 3
    try:
        with pytest.raises(Exception, match="bflux_chain"):
            bk.rate(U0)
    finally:
        bk.P.bflux_chain = 0
