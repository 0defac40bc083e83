"""The CPU oracle against the reference's golden vectors.

Pins every oracle function that the device parity tests lean on: the edge
and element passes for all three fluxes and all four boundary kinds, the
folded rate, the inverse mass, the time step, the weighted norm, the
positivity limiter (outputs and activation flags), the superbee limiter and
whole SSP-RK3/Euler steps.  Tolerances: residuals 1e-12 relative (max norm),
everything else 1e-13, flags exact.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import build_case, golden, nz_module, rel_err

from oracle.oracle import CpuOracle

FLUXES = ("roe", "hllc", "ausmup")


@pytest.mark.parametrize("flux", FLUXES)
def test_residual_cfg1(flux):
    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    orc = CpuOracle(ops, flow.replace(flux=flux))
    for st in ("U0", "Ur"):
        R, w = orc.residual(g[st])
        assert rel_err(R, g[f"R_{flux}_{st}"]) <= 1e-12
        assert rel_err(w, g[f"wave_{flux}_{st}"]) <= 1e-13


@pytest.mark.parametrize("name,key", [("quad_q2_p2", "U0"), ("tri_p2_q2_backflow", "U"),
                                      ("synth_p3", "U"), ("synth_p1", "U")])
@pytest.mark.parametrize("flux", FLUXES)
def test_residual_other_cases(name, key, flux):
    g = golden(name)
    ops, flow = build_case(name)
    R, w = CpuOracle(ops, flow.replace(flux=flux)).residual(g[key])
    suffix = f"_{key}" if name == "quad_q2_p2" else ""
    assert rel_err(R, g[f"R_{flux}{suffix}"]) <= 1e-12
    assert rel_err(w, g[f"wave_{flux}{suffix}"]) <= 1e-13


def test_rate_dt_norm_cfg1():
    cf = nz_module("config")
    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    orc = CpuOracle(ops, flow)
    F, w = orc.rate(g["U0"])
    assert rel_err(F, g["rate_roe_U0"]) <= 1e-12
    R, _ = orc.residual(g["U0"])
    assert rel_err(-orc.inverse_mass(R), g["rate_roe_U0"]) <= 1e-12
    dt = orc.time_step(w, cf.step_coefficient(1, None, "ssprk3"))
    assert rel_err(dt, g["dt_U0"]) <= 1e-13
    assert abs(orc.norm(F) / float(g["norm_rate_roe_U0"]) - 1) <= 1e-13


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2"])
def test_positivity_limiter(name):
    g = golden(name)
    ops, flow = build_case(name)
    out, flags, (n_scaled, n_rep) = CpuOracle(ops, flow).limit(g["Us"])
    assert np.array_equal(flags, g["Us_flags"])
    assert n_scaled == int(((g["Us_flags"] & 2) > 0).sum())
    assert n_rep == int(((g["Us_flags"] & 1) > 0).sum())
    assert rel_err(out, g["Us_positivity"]) <= 1e-13
    # the screen never changes the answer
    out2, flags2, _ = CpuOracle(ops, flow, screen=False).limit(g["Us"])
    assert np.array_equal(out, out2) and np.array_equal(flags, flags2)


def test_limiter_inactive_and_conservative():
    """Exactly inactive on an admissible smooth state; conserves cell means."""
    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    orc = CpuOracle(ops, flow)
    out, flags, cnt = orc.limit(g["U0"])
    assert np.array_equal(out, g["U0"]) and cnt == (0, 0) and not flags.any()
    out, flags, _ = orc.limit(g["Us"])
    mw = ops.mean_weights
    mean_in = np.einsum("ei,eis->es", mw, g["Us"])
    mean_out = np.einsum("ei,eis->es", mw, out)
    ok = (flags & 1) == 0
    assert np.abs(mean_out[ok] - mean_in[ok]).max() <= 1e-13 * np.abs(mean_in).max()


def test_superbee():
    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    orc = CpuOracle(ops, flow)
    out, _, _ = orc.limit(g["Ub"], "superbee", 0.0)
    assert rel_err(out, g["Ub_apply_superbee"]) <= 1e-13


@pytest.mark.parametrize("scheme,steps,key", [("ssprk3", 1, "ssprk3_1step"),
                                              ("ssprk3", 10, "ssprk3_10step"),
                                              ("euler", 1, "euler_1step")])
def test_steps_cfg1(scheme, steps, key):
    cf = nz_module("config")
    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    U, res = CpuOracle(ops, flow).run(g["U0"], steps, scheme,
                                      cf.step_coefficient(1, None, scheme))
    assert rel_err(U, g[f"{key}_U"]) <= 1e-12
    assert abs(res / float(g[f"{key}_res"]) - 1) <= 1e-13


def test_steps_quad_and_p0():
    cf = nz_module("config")
    g = golden("quad_q2_p2")
    ops, flow = build_case("quad_q2_p2")
    U, res = CpuOracle(ops, flow).run(g["U0"], 1, "ssprk3", cf.step_coefficient(2, None))
    assert rel_err(U, g["ssprk3_1step_U"]) <= 1e-12
    g = golden("tri_p0")
    ops, flow = build_case("tri_p0")
    U, res = CpuOracle(ops, flow).run(g["U0"], 5, "euler", cf.step_coefficient(0, None, "euler"))
    assert str(g["scheme"]) == "euler"
    assert rel_err(U, g["5step_U"]) <= 1e-12
    assert abs(res / float(g["5step_res"]) - 1) <= 1e-12


def test_planted_sign_error_is_caught():
    """A dropped sign in the lift would be caught by the goldens."""
    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    orc = CpuOracle(ops, flow)
    orc.face_sign = -orc.face_sign  # plant the mistake in the oracle, on the CPU
    R, _ = orc.residual(g["U0"])
    assert rel_err(R, g["R_roe_U0"]) > 1e-3


def test_standalone_limiter_matches_oracle_limit():
    """The operator-layout-free limiter entry point agrees with CpuOracle.limit."""
    from oracle.oracle import CpuOracle, positivity_limit

    ops, flow = build_case("quad_q2_p2", dense=True)
    g = golden("quad_q2_p2")
    a_U, a_f, a_c = CpuOracle(ops, flow).limit(g["Us"], "positivity")
    b_U, b_f, b_c = positivity_limit(g["Us"], ops, flow)
    assert np.array_equal(a_U, b_U) and np.array_equal(a_f, b_f) and a_c == b_c
    assert np.array_equal(b_f, g["Us_flags"])
