"""The smooth nozzle end to end on the device (AR 2.5, P1 triangles,
refine 1, Roe, positivity limiter, SSP-RK3) marched from the reference's
quasi-1D initial state to a scaled residual of 1e-10.

Bars (per result): thrust, mass flow in/out and exit Mach within 1e-10
relative of the reference's converged values, and the same stop iteration.
The limiter is inactive along this path, so both runs see the same sequence of
smooth states and stop at the same convergence check.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import build_case, golden, nz_module, rel_err

pytestmark = pytest.mark.gpu


def test_converged_smooth_nozzle():
    cf = nz_module("config")
    mr = nz_module("march")
    ob = nz_module("observables")
    g = golden("cfg1_converged")
    U0 = golden("cfg1")["U0"]
    ops, flow = build_case("cfg1", dense=False)
    opts = cf.SolverOptions(tolerance=1e-10, print_interval=0)
    res = mr.solve_steady(ops, flow, opts, U0=U0, backend="hip")
    assert res.converged, res.message
    assert res.iterations == int(g["iterations"])
    perf = ob.performance(res.U, ops, flow)
    errs = {
        "thrust": abs(perf.thrust / float(g["perf_thrust"]) - 1),
        "mass_flow_in": abs(perf.mass_flow_in / float(g["perf_mass_flow_in"]) - 1),
        "mass_flow_out": abs(perf.mass_flow_out / float(g["perf_mass_flow_out"]) - 1),
        "exit_mach": abs(perf.exit_mach_area_averaged / float(g["perf_exit_mach_area_averaged"]) - 1),
        "exit_pressure": abs(perf.exit_pressure_ratio / float(g["perf_exit_pressure_ratio"]) - 1),
        "thrust_wall": abs(perf.thrust_wall / float(g["perf_thrust_wall"]) - 1),
    }
    print("converged in", res.iterations, "iterations;", res.history.wall_time, "s; errors", errs)
    for k, v in errs.items():
        assert v <= 1e-10, (k, v)
    assert rel_err(res.U, g["U"]) <= 1e-9
    assert res.mean_repairs == int(g["mean_repairs"])


@pytest.mark.parametrize("limiter", ["positivity", "superbee"])
def test_graph_chunks_match_eager(limiter):
    """``run`` replays a captured HIP graph for a chunk length it has seen
    before.  Every replay must give the eager chunk's state and residual bit
    for bit, with other eager work on the same backend between replays."""
    cf = nz_module("config")
    be = nz_module("backend")
    ops, flow = build_case("cfg1", dense=False)
    U0 = golden("cfg1")["U0"]
    opts = cf.SolverOptions(limiter=limiter)
    eager = be.HipBackend(ops, flow, opts)
    eager.use_graphs = False
    graph = be.HipBackend(ops, flow, opts)
    Ue, Ug = U0, U0
    for k in range(5):
        Ue, re = eager.run(Ue, 50, "ssprk3")
        Ug, rg = graph.run(Ug, 50, "ssprk3")
        assert np.array_equal(Ue, Ug) and re == rg, k
        assert len(graph._graphs) == (0 if k == 0 else 1)
        graph.residual(Ug)  # eager work between replays
        graph.run(Ug, 1, "ssprk3")
    assert graph.n_limit_calls == eager.n_limit_calls + 5 * 3
