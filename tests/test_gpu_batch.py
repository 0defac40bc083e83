"""The batched multi-design march against one-at-a-time solves on the device.

Designs of three families (tri P1, tri P2, quad P1) are marched together with
per-design convergence masking and HIP-graph replay of each family's chunk;
each design must stop at the same iteration with the same bits as the same
design marched alone through the ordinary backend (eager launches).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import nz_module

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sweep():
    batch = nz_module("batch")
    cf = nz_module("config")
    designs = [
        batch.Design(area_ratio=2.0, kind="tri", order=1, refine=0),
        batch.Design(area_ratio=3.2, kind="tri", order=1, refine=0),
        batch.Design(area_ratio=2.5, kind="tri", order=2, refine=0),
        batch.Design(area_ratio=1.8, kind="quad", order=1, refine=0),
        batch.Design(area_ratio=3.6, kind="quad", order=1, refine=1),
    ]
    flow = cf.FlowConditions()
    opts = cf.SolverOptions(tolerance=1e-8, max_iterations=6000)
    run = batch.BatchedSteady(designs, flow, opts)
    print(f"batched sweep: {len(designs)} designs, {run.n_elements} elements, "
          f"{len(run.families)} families", flush=True)
    return run, run.solve(use_graph=True), flow, opts


def test_batched_designs_match_lone_solves(sweep):
    run, results, flow, opts = sweep
    march = nz_module("march")
    for g, res in enumerate(results):
        lone = march.solve_steady(run.ops_list[g], flow, opts, U0=run.U0_list[g])
        print(f"design {g}: batch {res.iterations} it conv={res.converged}, "
              f"lone {lone.iterations} it conv={lone.converged}", flush=True)
        assert res.converged == lone.converged
        assert res.iterations == lone.iterations
        assert res.residual == lone.residual
        assert np.array_equal(res.U, lone.U)


def test_batched_designs_converged_and_masked(sweep):
    run, results, flow, opts = sweep
    iters = [r.iterations for r in results]
    assert all(r.converged for r in results), [r.message for r in results]
    # designs stop at different iterations; the later ones kept marching alone
    assert len(set(iters)) > 1
    obs = nz_module("observables")
    for r in results:
        perf = obs.performance(r.U, r.operators, r.flow)
        assert np.isfinite(perf.thrust) and perf.thrust > 0.0
        assert perf.mass_flow_out == pytest.approx(perf.mass_flow_in, rel=1e-2)


def test_eager_and_graph_replay_agree(sweep):
    """The same sweep without graph capture gives the same bits."""
    run, results, flow, opts = sweep
    batch = nz_module("batch")
    eager = batch.BatchedSteady(run.designs, flow, opts, ops_list=run.ops_list,
                                U0_list=run.U0_list).solve(use_graph=False)
    for a, b in zip(results, eager):
        assert a.iterations == b.iterations
        assert np.array_equal(a.U, b.U)


def test_batched_gradients_match_lone_adjoints(sweep):
    """Thrust gradients w.r.t. the area ratio from the batched adjoint march
    agree with each design's own adjoint solve (DifferentiableCase)."""
    run, results, flow, opts = sweep
    batch, sens = nz_module("batch"), nz_module("sensitivity")
    grads = batch.batched_gradients(results, flow, opts, tolerance=1e-8)
    for g, (res, gr) in enumerate(zip(results, grads)):
        assert gr is not None and gr.converged
        case = sens.DifferentiableCase(geometry=res.design.geometry(), flow=flow,
                                       order=res.design.order, refine=res.design.refine,
                                       kind=res.design.kind, options=opts)
        J, grad, fwd, info = case.value_and_gradient("thrust", names=("area_ratio",),
                                                     U0=run.U0_list[g], adjoint_tolerance=1e-8)
        print(f"design {g}: batch dJ/dAR {gr.gradient['area_ratio']:.10e} "
              f"({gr.adjoint_iterations} it), lone {grad['area_ratio']:.10e} "
              f"({info['adjoint_iterations']} it)", flush=True)
        assert J == gr.objective
        assert gr.adjoint_iterations == info["adjoint_iterations"]
        assert gr.gradient["area_ratio"] == pytest.approx(grad["area_ratio"], rel=1e-10)
