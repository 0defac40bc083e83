"""The batched parameter sweep with the reference's table and CSV
(``src/sweep.py:300-446``): every point's metrics against a lone solve of
that point, and the warm start's saving."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import nz_module

pytestmark = pytest.mark.gpu


def test_sweep_table_matches_lone_solves(tmp_path):
    batch, cf, mr, out = (nz_module("batch"), nz_module("config"), nz_module("march"),
                          nz_module("outputs"))
    ob, ini = nz_module("observables"), nz_module("initial")
    opts = cf.SolverOptions(tolerance=1e-8, print_interval=0)
    flow = cf.FlowConditions()
    res = batch.sweep(area_ratio=np.linspace(2.0, 3.0, 3), refine=1, order=1, options=opts)
    assert res.all_converged and res.shape == (3,)
    ops_list = batch.build_device_designs(
        [batch.Design(area_ratio=a, refine=1) for a in (2.0, 2.5, 3.0)])
    for i, a in enumerate((2.0, 2.5, 3.0)):
        geom = batch.Design(area_ratio=a).geometry()
        bk = nz_module("backend").HipBackend(ops_list[i], flow, opts)
        U0 = ini.quasi1d_initial(ops_list[i], flow, geom, backend=bk)
        lone = mr.solve_steady(ops_list[i], flow, opts, U0=U0)
        perf = ob.performance(lone.U, ops_list[i], flow, geom=geom)
        assert res["iterations"][i] == lone.iterations
        for m, f in (("thrust", "thrust"), ("thrust_coefficient", "thrust_coefficient"),
                     ("exit_mach", "exit_mach_area_averaged"), ("entropy_error", "entropy_error"),
                     ("ideal_thrust", "ideal_thrust")):
            assert abs(res[m][i] / getattr(perf, f) - 1) <= 1e-12, (m, res[m][i], getattr(perf, f))
        assert abs(res["min_density"][i] / lone.min_density - 1) <= 1e-12
    path = tmp_path / "sweep.csv"
    res.to_csv(str(path))
    cols = out.read_csv(str(path))
    assert [float(v) for v in cols["thrust"]] == list(res["thrust"])


def test_sweep_warm_start_saves_iterations():
    batch, cf = nz_module("batch"), nz_module("config")
    opts = cf.SolverOptions(tolerance=1e-8, print_interval=0)
    grid = np.linspace(2.0, 2.6, 7)
    cold = batch.sweep(area_ratio=grid, refine=1, options=opts)
    warm = batch.sweep(area_ratio=grid, refine=1, options=opts, warm_start=True)
    assert cold.all_converged and warm.all_converged
    print("iterations cold", cold["iterations"], "warm", warm["iterations"])
    assert warm["iterations"][1::2].sum() < 0.8 * cold["iterations"][1::2].sum()
    # the same steady states, to the solver tolerance
    assert np.allclose(warm["thrust"], cold["thrust"], rtol=1e-6)
