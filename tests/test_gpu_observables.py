"""Observables and adjoint objectives on the device (``nz_boundary_integrals``,
``nz_performance_integrals``, ``nz_objective``) against the host restatement
and the reference's figures.

Bars: every performance field within 1e-12 relative of the host's (which is
pinned to the reference's converged-state fixture); each objective's value,
state partial and edge-geometry partial within 1e-12 of the host's, on a
straight-edged triangle mesh and a curved (Q2) quadrilateral mesh; the
adjoint gradient of the averaged objectives and of the thrust coefficient
against central differences of the converged objective at 1e-4 (the
reference test's bar, step 1e-4 relative).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import build_case, golden, nz_module, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("thrust", "thrust_wall", "mass_flow_in", "mass_flow_out", "exit_mach_area_averaged",
          "exit_pressure_ratio", "thrust_imbalance", "thrust_coefficient", "mass_imbalance",
          "entropy_error", "exit_temperature_ratio", "specific_thrust", "specific_thrust_ratio",
          "discharge_coefficient", "ideal_thrust", "thrust_efficiency")


def _setup(name):
    cf, be, mesh = nz_module("config"), nz_module("backend"), nz_module("mesh")
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    from conftest import CASES

    mk = CASES[name][0]
    geom = mesh.NozzleGeometry(contour=mk["contour"], area_ratio=mk["area_ratio"])
    return g, ops, flow, geom, be.HipBackend(ops, flow, cf.SolverOptions())


@pytest.mark.parametrize("name,key", [("cfg1", None), ("quad_q2_p2", "U0")])
def test_device_performance_every_field(name, key):
    ob = nz_module("observables")
    g, ops, flow, geom, bk = _setup(name)
    U = golden("cfg1_converged")["U"] if key is None else g[key]
    dev = bk.performance(U, geom=geom)
    host = ob.performance(U, ops, flow, geom=geom)
    for k in FIELDS:
        a, b = getattr(dev, k), getattr(host, k)
        err = abs(a - b) / max(abs(b), 1e-300)
        print(name, k, a, err)
        assert err <= 1e-12 or (k.endswith("imbalance") and abs(a - b) <= 1e-14), k
    if key is None:
        gc = golden("cfg1_converged")
        for k in ("thrust_coefficient", "discharge_coefficient", "exit_temperature_ratio"):
            assert abs(getattr(dev, k) / float(gc[f"perf_{k}"]) - 1) <= 1e-12, k


@pytest.mark.parametrize("objective", ["thrust", "thrust_coefficient", "mass_flow", "exit_mach",
                                       "exit_pressure"])
@pytest.mark.parametrize("name,key", [("cfg1", None), ("quad_q2_p2", "U0")])
def test_device_objective_matches_host(objective, name, key):
    sens = nz_module("sensitivity")
    g, ops, flow, geom, bk = _setup(name)
    U = golden("cfg1_converged")["U"] if key is None else g[key]
    J, dU, dG = bk.objective(U, objective, geom=geom)
    Jh, dUh, dGh = sens.objective_and_partials(U, ops, flow, objective, geom=geom)
    dG = dG.cpu().numpy()
    print(name, objective, J, Jh, rel_err(dU.cpu().numpy(), dUh), rel_err(dG, dGh))
    assert abs(J / Jh - 1) <= 1e-12
    assert rel_err(dU.cpu().numpy(), dUh) <= 1e-12
    if objective in ("exit_mach", "exit_pressure"):
        # d/d|dX/dsigma| of an average is (Q - J) w / area: on a nearly
        # uniform exit Q - J cancels, so the bar is scaled by Q w / area
        ob = nz_module("observables")
        tr = ob.boundary_trace(U, ops, nz_module("mesh").BoundaryTag.OUTFLOW)
        scale = abs(J) * float(ops.ref.w_face.max()) / float(tr.weight.sum())
        assert np.abs(dG - dGh).max() <= 1e-12 * max(scale, np.abs(dGh).max())
    else:
        assert rel_err(dG, dGh) <= 1e-12


@pytest.mark.parametrize("objective,param", [("exit_mach", "area_ratio"),
                                             ("exit_pressure", "throat_x"),
                                             ("thrust_coefficient", "throat_half_height")])
def test_adjoint_gradient_of_other_objectives(objective, param):
    """The captured sensitivity case of the reference (smooth, p = 1,
    refine 0, pb 0.15, tolerance 1e-11): the adjoint gradient of an averaged
    objective and of the thrust coefficient against central differences of
    the converged objective (step 1e-4 relative, as the reference test)."""
    sens, cf, mesh, ini = (nz_module("sensitivity"), nz_module("config"), nz_module("mesh"),
                           nz_module("initial"))
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    case = sens.DifferentiableCase(geometry=geom, order=1, refine=0,
                                   options=cf.SolverOptions(tolerance=1e-11, print_interval=0,
                                                            max_iterations=200_000))
    U0 = ini.quasi1d_initial(case.operators(), case.flow, geom)
    J, grad, fwd, _ = case.value_and_gradient(objective, names=(param,), U0=U0,
                                              adjoint_tolerance=1e-10)
    base = case.default_params()
    h = 1e-4 * max(abs(base[param]), 1e-3)
    vals = []
    for sgn in (1.0, -1.0):
        p = {**base, param: base[param] + sgn * h}
        f = case.solve_forward(p, U0=fwd.U)
        assert f.converged, f.message
        gp = mesh.NozzleGeometry(contour="smooth", area_ratio=p["area_ratio"], throat_x=p["throat_x"],
                                 throat_half_height=p["throat_half_height"])
        vals.append(sens.objective_and_partials(f.U, f.operators, case.flow_at(p), objective,
                                                geom=gp)[0])
    fd = (vals[0] - vals[1]) / (2 * h)
    rel = abs(grad[param] - fd) / max(abs(fd), abs(grad[param]))
    print(f"d {objective} / d {param}: adjoint {grad[param]:.9e}, central difference {fd:.9e}, "
          f"rel {rel:.1e}")
    assert rel < 1e-4
