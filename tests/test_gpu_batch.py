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


def test_serialized_families_give_the_same_bits(sweep):
    """Families queued on one stream instead of side by side on their own."""
    run, results, flow, opts = sweep
    batch = nz_module("batch")
    serial = batch.BatchedSteady(run.designs, flow, opts, ops_list=run.ops_list,
                                 U0_list=run.U0_list).solve(concurrent=False)
    for a, b in zip(results, serial):
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


# ------------------------------------------------------------ full sweep
@pytest.fixture(scope="module")
def sweep256():
    """The benchmarked sweep: 256 designs from default_rng(1) (area ratio
    U[1.5, 4], refine and order in {1, 2, 3}, triangles and quadrilaterals),
    Roe, positivity limiter, tolerance 1e-10, at most 50,000 iterations."""
    batch, cf = nz_module("batch"), nz_module("config")
    designs = batch.design_sweep(256, 1)
    flow = cf.FlowConditions()
    opts = cf.SolverOptions(tolerance=1e-10, max_iterations=50000, print_interval=0)
    run = batch.BatchedSteady(designs, flow, opts)
    results = run.solve(use_graph=True)
    print(f"256 designs, {run.n_elements} elements, batch {run.wall_s:.1f} s, "
          f"{sum(r.converged for r in results)} converged", flush=True)
    return run, results, flow, opts


def _sample(results) -> list[int]:
    """Eight designs covering the families the sweep holds, the hardest first:
    tri p3, quad p2 and quad p3 at refine 3, then one of every other
    (kind, order) and two more at refine 3."""
    picks: list[int] = []

    def first(pred):
        for i, r in enumerate(results):
            if i not in picks and pred(r.design):
                picks.append(i)
                return

    for kind, order in (("tri", 3), ("quad", 2), ("quad", 3)):
        first(lambda d, k=kind, o=order: d.kind == k and d.order == o and d.refine == 3)
    for kind, order in (("tri", 1), ("tri", 2), ("quad", 1)):
        first(lambda d, k=kind, o=order: d.kind == k and d.order == o)
    first(lambda d: d.refine == 3)
    first(lambda d: d.refine == 3)
    return picks


def test_sweep256_sampled_designs_bit_identical_to_lone(sweep256):
    run, results, flow, opts = sweep256
    march = nz_module("march")
    picks = _sample(results)
    assert len(picks) == 8
    for g in picks:
        res = results[g]
        d = res.design
        lone = march.solve_steady(run.ops_list[g], flow, opts, U0=run.U0_list[g])
        print(f"design {g} ({d.kind} p{d.order} refine {d.refine}, AR {d.area_ratio:.3f}): "
              f"batch {res.iterations} it conv={res.converged}, lone {lone.iterations} it",
              flush=True)
        assert (res.converged, res.iterations) == (lone.converged, lone.iterations)
        assert res.residual == lone.residual or (np.isnan(res.residual) and np.isnan(lone.residual))
        assert np.array_equal(res.U, lone.U, equal_nan=True)


def test_sweep256_batched_gradients_equal_lone_adjoints(sweep256):
    run, results, flow, opts = sweep256
    batch, sens = nz_module("batch"), nz_module("sensitivity")
    grads = batch.batched_gradients(results, flow, opts, tolerance=1e-8)
    n_conv = sum(r.converged for r in results)
    assert sum(1 for gr in grads if gr is not None and gr.converged) == n_conv
    checked = 0
    for g in _sample(results):
        res, gr = results[g], grads[g]
        if not res.converged:
            assert gr is None
            continue
        d = res.design
        case = sens.DifferentiableCase(geometry=d.geometry(), flow=flow, order=d.order,
                                       refine=d.refine, kind=d.kind, options=opts)
        J, grad, fwd, info = case.value_and_gradient("thrust", names=("area_ratio",),
                                                     U0=run.U0_list[g], adjoint_tolerance=1e-8)
        print(f"design {g}: batch dJ/dAR {gr.gradient['area_ratio']:.12e} "
              f"({gr.adjoint_iterations} it), lone {grad['area_ratio']:.12e} "
              f"({info['adjoint_iterations']} it)", flush=True)
        assert J == gr.objective and gr.adjoint_iterations == info["adjoint_iterations"]
        assert gr.gradient["area_ratio"] == grad["area_ratio"]
        checked += 1
    assert checked >= 4


def test_device_setup_batch(sweep):
    """The same designs with the grids' connectivity built once on the device
    and shared, operators and initial states built on the device: each design
    matches its lone solve on the same operators bit for bit, and the
    host-built batch to round-off (same stop iterations, observables within
    1e-10)."""
    run_h, results_h, flow, opts = sweep
    batch, march, obs = nz_module("batch"), nz_module("march"), nz_module("observables")
    run = batch.BatchedSteady(run_h.designs, flow, opts, device_setup=True)
    assert run.n_topologies == 3  # tri r0, quad r0, quad r1 grids
    results = run.solve(use_graph=True)
    for g, (res, rh) in enumerate(zip(results, results_h)):
        lone = march.solve_steady(run.ops_list[g], flow, opts, U0=run.U0_list[g])
        assert (res.iterations, res.converged) == (lone.iterations, lone.converged)
        assert np.array_equal(res.U, lone.U)
        assert (res.iterations, res.converged) == (rh.iterations, rh.converged)
        p, ph = obs.performance(res.U, res.operators, flow), obs.performance(rh.U, rh.operators, flow)
        err = abs(p.thrust / ph.thrust - 1)
        print(f"design {g}: device setup {res.iterations} it, thrust rel diff {err:.1e}", flush=True)
        assert err <= 1e-10
