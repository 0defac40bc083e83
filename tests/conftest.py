"""Shared test helpers.

``gpu``-marked tests need an MI355X and the built HIP library; everything
else runs on the CPU.  The package directory ``2d_nozzle_amd`` is imported
under the alias ``nz`` (its name is not a Python identifier).
"""

from __future__ import annotations

import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running end-to-end solves")


def nz_module(name: str = ""):
    return importlib.import_module("2d_nozzle_amd" + (f".{name}" if name else ""))


def golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def rel_err(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ---------------------------------------------------------------- the golden cases
CASES = {
    # name: (mesh builder kwargs, order, flow kwargs)
    "cfg1": (dict(contour="smooth", area_ratio=2.5, kind="tri", refine=1, q=1), 1, {}),
    "quad_q2_p2": (dict(contour="bell", area_ratio=2.5, kind="quad", refine=0, q=2), 2,
                   dict(back_pressure_ratio=0.7, allow_shock_in_nozzle=True, flux="hllc")),
    "tri_p0": (dict(contour="smooth", area_ratio=2.5, kind="tri", refine=0, q=1), 0, {}),
    "tri_p2_q2_backflow": (dict(contour="smooth", area_ratio=2.5, kind="tri", refine=0, q=2), 2,
                           dict(back_pressure_ratio=0.9, backflow=True, hllc_low_mach=0.3)),
}


def build_case(name: str, dense: bool = True):
    """(ops, flow) of a golden case, built with this package's host code."""
    mesh = nz_module("mesh")
    op = nz_module("operators")
    cf = nz_module("config")
    if name.startswith("synth_"):
        order = int(name[-1])
        ncx = {3: 32, 1: 12}[order]
        topo, coords = mesh.synthetic_channel(ncx, ncx // 2)
        return op.build_operators(topo, coords, order, dense=dense), cf.FlowConditions()
    mk, order, fk = CASES[name]
    geom = mesh.NozzleGeometry(contour=mk["contour"], area_ratio=mk["area_ratio"])
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=mk["kind"], refine=mk["refine"],
                                             geometry_order=mk["q"])
    return op.build_operators(topo, coords, order, dense=dense), cf.FlowConditions(**fk)


@pytest.fixture(scope="session")
def nz():
    return nz_module()
