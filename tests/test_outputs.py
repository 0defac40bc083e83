"""Output formats (CPU): the solution .npz of the reference's ``solve --save``
(``src/cli.py:166-174``) and the sweep table (``src/sweep.py:68-145``)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import nz_module


def test_save_solution_keys(tmp_path):
    out, mesh = nz_module("outputs"), nz_module("mesh")

    class R:
        U = np.arange(24.0).reshape(2, 3, 4)
        converged = True
        residual = 3.5e-11

    geom = mesh.NozzleGeometry(contour="bezier9", area_ratio=2.2)
    path = tmp_path / "sol.npz"
    out.save_solution(str(path), R(), geom)
    with np.load(path, allow_pickle=False) as d:
        assert np.array_equal(d["U"], R.U) and bool(d["converged"]) and d["residual"] == 3.5e-11
        assert d["geom_contour"] == "bezier9" and d["geom_area_ratio"] == 2.2
        assert "geom_theta_initial_deg" not in d.files  # None fields are left out


def _table():
    out = nz_module("outputs")
    pts = [{"area_ratio": a} for a in (2.0, 2.5, 3.0)]
    vals = {m: np.array([1.0 / 3.0, math.pi, np.nan]) for m in out.METRICS}
    return out.SweepResult(parameters=("area_ratio",), points=pts, values=vals,
                           converged=np.array([True, True, False]),
                           regimes=["overexpanded"] * 3, messages=["", "", "stalled"],
                           shape=(3,))


def test_sweep_result_indexing_and_failures():
    t = _table()
    assert t.n_points == 3 and not t.all_converged
    assert np.array_equal(t["area_ratio"], [2.0, 2.5, 3.0])
    assert t["thrust"][1] == math.pi and t.reshape("thrust").shape == (3,)
    assert t.failures() == [({"area_ratio": 3.0}, "stalled")]
    with pytest.raises(KeyError, match="neither a swept parameter"):
        t["nope"]
    d = t.to_dict()
    assert set(d) == {"area_ratio", *nz_module("outputs").METRICS, "converged"}


def test_sweep_csv_round_trip(tmp_path):
    out = nz_module("outputs")
    t = _table()
    path = tmp_path / "sweep.csv"
    t.to_csv(str(path))
    cols = out.read_csv(str(path))
    assert list(cols)[:1] == ["area_ratio"] and list(cols)[-2:] == ["converged", "regime"]
    assert list(cols)[1:-2] == list(out.METRICS)
    assert [float(v) for v in cols["thrust"][:2]] == [1.0 / 3.0, math.pi]  # exact repr
    assert math.isnan(float(cols["thrust"][2]))
    assert cols["converged"] == ["True", "True", "False"]
    text = t.table()
    assert "area_ratio" in text and "NO" in text.splitlines()[-1]


def test_sweep_serpentine_and_argument_checks():
    batch = nz_module("batch")
    assert batch._serpentine((2, 3)) == [(0, 0), (0, 1), (0, 2), (1, 2), (1, 1), (1, 0)]
    assert batch._serpentine((3,)) == [(0,), (1,), (2,)]
    with pytest.raises(NotImplementedError, match="back_pressure_ratio"):
        batch.sweep(back_pressure_ratio=[0.1, 0.2])
    with pytest.raises(ValueError, match="at least one iterable"):
        batch.sweep()
