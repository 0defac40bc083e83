"""Why the quadrilateral p = 2 and p = 3 designs of the 256-design sweep do
not converge: the reference's own scheme does not converge them.

The CPU oracle (the reference's compiled kernels restated, pinned to the
reference's goldens) marches the first quad p = 3 and quad p = 2 designs of
``design_sweep(256, 1)`` (both refine 1, 280 elements) from their quasi-1D
states with the sweep's settings (Roe, positivity limiter, SSP-RK3 at the
recommended cfl, tolerance 1e-10).  The p = 3 design's residual jumps about
700-fold within the first 50 steps and stays there; the p = 2 design falls to
about 1e-4 (scaled) and then wanders between 4e-5 and 5e-4 without
converging.  The device march does the same (the benchmark reports these
designs stopped by the reference's own stall, divergence and budget rules).
"""

from __future__ import annotations

import numpy as np

from conftest import nz_module


def _oracle_march(design, steps: int, check: int = 50):
    from oracle.oracle import CpuOracle

    mesh, opm, cf, ini = (nz_module("mesh"), nz_module("operators"), nz_module("config"),
                          nz_module("initial"))
    geom = design.geometry()
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=design.kind, nx=design.nx, nr=design.nr,
                                             refine=design.refine)
    ops = opm.build_operators(topo, coords, design.order, dense=True)
    flow = cf.FlowConditions()
    U = ini.quasi1d_initial(ops, flow, geom)
    orc = CpuOracle(ops, flow)
    fac = cf.step_coefficient(design.order, None, "ssprk3")
    scale = cf.residual_scale(flow, 1.0)
    hist = []
    U, r = orc.run(U, 1, "ssprk3", fac)
    hist.append(r / scale)
    for _ in range(0, steps - 1, check):
        U, r = orc.run(U, check, "ssprk3", fac)
        hist.append(r / scale)
    return np.asarray(hist)


def _first(kind: str, order: int):
    batch = nz_module("batch")
    ds = batch.design_sweep(256, 1)
    picks = [(d.refine, i) for i, d in enumerate(ds) if d.kind == kind and d.order == order]
    return ds[min(picks)[1]]


def test_quad_p3_design_blows_up_under_the_reference_scheme():
    d = _first("quad", 3)
    assert (d.refine, d.order) == (1, 3)
    h = _oracle_march(d, 401)
    assert np.isfinite(h).all()
    assert h[1] > 300 * h[0]                    # a jump of hundreds at the first check
    assert h[1:].min() > 100 * h[0]             # and no recovery
    assert h.min() > 1e6 * 1e-10                # nowhere near the tolerance


def test_quad_p2_design_wanders_without_converging():
    d = _first("quad", 2)
    assert (d.refine, d.order) == (1, 2)
    h = _oracle_march(d, 15001)
    late = h[100:]                              # after 5,000 iterations
    assert late.min() > 1e-6                    # four orders above the 1e-10 tolerance
    assert late.max() > 5 * late.min()          # not a monotone decay
    assert h[-1] > 2 * late.min()               # it has climbed back up by the end
