"""Initial states and quasi-1D theory (host side).

The projected quasi-1D states are checked against the reference's own initial
states held in the fixtures (smooth tri P1, bell quad Q2 P2 with a normal
shock, tri P0).  The theory is checked against closed forms and the known
answers of the reference's quasi-1D tests (src tests/test_quasi1d.py:67-72).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import CASES, build_case, golden, nz_module, rel_err

G = 1.4


@pytest.fixture(scope="module")
def ini():
    return nz_module("initial")


@pytest.fixture(scope="module")
def mesh():
    return nz_module("mesh")


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2", "tri_p0"])
def test_quasi1d_initial_matches_reference_state(ini, mesh, name):
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    mk = CASES[name][0]
    geom = mesh.NozzleGeometry(contour=mk["contour"], area_ratio=mk["area_ratio"])
    U0 = ini.initial_state(ops, flow, geom, "quasi1d")
    assert U0.shape == g["U0"].shape
    assert rel_err(U0, g["U0"]) < 1e-13


def test_shocked_case_has_a_shock_in_the_diverging_section(ini, mesh):
    geom = mesh.NozzleGeometry(contour="bell", area_ratio=2.5)
    flow = nz_module("config").FlowConditions(back_pressure_ratio=0.7, allow_shock_in_nozzle=True)
    sol = ini.solve_quasi1d(geom, flow)
    assert sol.regime is ini.Regime.SHOCK_IN_NOZZLE
    assert geom.throat_location() < sol.shock_x < geom.length
    assert sol.exit_pressure == pytest.approx(0.7, rel=1e-6)
    # Rankine-Hugoniot across the shock
    m2 = float(ini.normal_shock_mach(sol.shock_mach, G))
    i = int(np.searchsorted(sol.x, sol.shock_x))
    assert sol.mach[i + 1] == pytest.approx(m2, rel=0.05)
    mdot = sol.density * sol.velocity * sol.area
    assert (mdot.max() - mdot.min()) / mdot.mean() < 1e-8


@pytest.mark.parametrize("mach", [0.05, 0.2, 0.5, 0.9, 1.4, 2.0, 3.0, 5.0])
def test_area_mach_inversion_round_trips(ini, mach):
    ar = float(ini.area_over_throat(mach, G))
    assert float(ini.mach_from_area(ar, G, supersonic=mach > 1.0)) == pytest.approx(mach, abs=1e-9)


def test_closed_forms(ini):
    assert float(ini.area_over_throat(1.0, G)) == pytest.approx(1.0, abs=1e-14)
    for m in (0.1, 0.5, 1.0, 2.5):
        r = float(ini.pressure_ratio(m, G))
        assert float(ini.mach_from_pressure_ratio(r, G)) == pytest.approx(m, abs=1e-10)
    # normal shock: M2 < 1, p2/p1 > 1, 0 < pt2/pt1 < 1; vanishing at M1 = 1
    for m1 in (1.2, 2.0, 5.0):
        assert float(ini.normal_shock_mach(m1, G)) < 1.0
        assert float(ini.normal_shock_static_pressure_ratio(m1, G)) > 1.0
        assert 0.0 < float(ini.normal_shock_total_pressure_ratio(m1, G)) < 1.0
    assert float(ini.normal_shock_total_pressure_ratio(1.0, G)) == pytest.approx(1.0, abs=1e-9)
    with pytest.raises(ValueError, match=r"A/A\* must be >= 1"):
        ini.mach_from_area(0.5, G)


def test_critical_ratios_reference_geometry(ini):
    c = ini.critical_ratios(2.50189, G)
    assert c.first == pytest.approx(0.9609, abs=1e-3)
    assert c.second == pytest.approx(0.4345, abs=1e-3)
    assert c.third == pytest.approx(0.06390, abs=1e-4)
    assert c.exit_mach_design == pytest.approx(2.4436, abs=1e-3)


@pytest.mark.parametrize("pb, expected", [(0.99, "subsonic"), (0.90, "shock-in-nozzle"),
                                          (0.50, "shock-in-nozzle"), (0.35, "overexpanded"),
                                          (0.15, "overexpanded"), (0.03, "underexpanded")])
def test_operating_regimes(ini, pb, expected):
    assert ini.operating_regime(2.50189, pb, G)[0].value == expected


def test_projection_is_exact_on_the_space(ini):
    """Projecting the point values of a discrete field returns the field."""
    ops, _ = build_case("quad_q2_p2", dense=False)
    rng = np.random.default_rng(3)
    U = rng.standard_normal((ops.n_elem, ops.ref.n_basis, 4))
    values = np.einsum("iq,eis->eqs", np.asarray(ops.ref.phi_vol), U)
    assert rel_err(ini.project(values, ops), U) < 1e-12


def test_uniform_initial_is_uniform(ini):
    ops, flow = build_case("cfg1", dense=False)
    U = ini.initial_state(ops, flow, None, "uniform")
    s = ini.uniform_state(flow, 0.95)
    assert rel_err(U, np.broadcast_to(s, U.shape)) < 1e-13
    with pytest.raises(ValueError, match="unknown initial_condition"):
        ini.initial_state(ops, flow, None, "bogus")


def test_change_order_raises_exactly(ini, mesh):
    """p-continuation: a P1 field is represented exactly at P2."""
    op = nz_module("operators")
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind="tri", refine=0)
    o1 = op.build_operators(topo, coords, 1)
    o2 = op.build_operators(topo, coords, 2)
    U1 = np.random.default_rng(4).standard_normal((o1.n_elem, o1.ref.n_basis, 4))
    U2 = ini.change_order(U1, o1, o2)
    back = ini.change_order(U2, o2, o1)
    assert rel_err(back, U1) < 1e-12
