"""Observables on the reference's converged state reproduce its figures (CPU)."""

from conftest import build_case, golden, nz_module


def test_performance_of_reference_state():
    ob = nz_module("observables")
    g = golden("cfg1_converged")
    ops, flow = build_case("cfg1", dense=False)
    perf = ob.performance(g["U"], ops, flow)
    assert abs(perf.thrust / float(g["perf_thrust"]) - 1) <= 1e-13
    assert abs(perf.thrust_wall / float(g["perf_thrust_wall"]) - 1) <= 1e-13
    assert abs(perf.mass_flow_in / float(g["perf_mass_flow_in"]) - 1) <= 1e-13
    assert abs(perf.mass_flow_out / float(g["perf_mass_flow_out"]) - 1) <= 1e-13
    assert abs(perf.exit_mach_area_averaged / float(g["perf_exit_mach_area_averaged"]) - 1) <= 1e-13
    assert abs(perf.exit_pressure_ratio / float(g["perf_exit_pressure_ratio"]) - 1) <= 1e-13


def test_all_performance_fields_of_reference_state():
    """The fields the reference's converged-state fixture holds beyond the
    boundary integrals (postprocess.py:193-255), with the geometry given."""
    ob, mesh = nz_module("observables"), nz_module("mesh")
    g = golden("cfg1_converged")
    ops, flow = build_case("cfg1", dense=False)
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    perf = ob.performance(g["U"], ops, flow, geom=geom)
    for k in ("thrust_coefficient", "discharge_coefficient", "exit_temperature_ratio"):
        assert abs(getattr(perf, k) / float(g[f"perf_{k}"]) - 1) <= 1e-13, k
    # identities between the fields
    assert perf.thrust_imbalance == abs(perf.thrust - perf.thrust_wall) / abs(perf.thrust)
    assert perf.specific_thrust == perf.thrust / perf.mass_flow_in
    assert perf.thrust_efficiency == perf.thrust / perf.ideal_thrust
    assert 0.0 < perf.entropy_error < 0.1 and 0.9 < perf.thrust_efficiency < 1.1


def test_entropy_error_closed_form():
    """A uniform state with p = c s_t rho^gamma has s / s_t = c at every
    point, so the entropy error is |c - 1| exactly (up to rounding)."""
    import numpy as np

    ob = nz_module("observables")
    ops, flow = build_case("quad_q2_p2", dense=False)
    g = flow.gamma
    s_t = flow.total_pressure ** (1.0 - g) * (flow.Rgas * flow.total_temperature) ** g
    rho, vx, vy, c = 0.7, 0.3, -0.1, 1.0375
    p = c * s_t * rho**g
    state = np.array([rho, rho * vx, rho * vy, p / (g - 1) + 0.5 * rho * (vx * vx + vy * vy)])
    U = np.broadcast_to(state, (ops.n_elem, ops.ref.n_basis, 4)).copy()
    area = float(np.asarray(ops.elem_area).sum())
    err = np.sqrt(ob.entropy_integral(U, ops, flow) / area)
    assert abs(err - (c - 1.0)) <= 1e-12


def test_march_logic_on_the_oracle():
    """The host march's stop rules, driven by the CPU oracle (no GPU)."""
    import numpy as np

    from oracle.oracle import CpuOracle

    cf = nz_module("config")
    mr = nz_module("march")
    ops, flow = build_case("cfg1")
    U0 = golden("cfg1")["U0"]

    class OracleBackend:
        scheme = "ssprk3"

        def __init__(self):
            self.o = CpuOracle(ops, flow)

        def run(self, U, n, scheme):
            return self.o.run(U, n, scheme, cf.step_coefficient(1, None, scheme))

        def limiter_activity(self):
            return float("nan")

    opts = cf.SolverOptions(tolerance=1e-10, max_iterations=101, print_interval=0)
    U, hist, conv, iters, res, res0, msg = mr.march(U0, OracleBackend(), opts, 1.0)
    assert not conv and iters == 101 and "max_iterations" in msg
    g = golden("cfg1_converged")
    assert np.allclose(hist.residual[:3], g["hist_residual"][:3], rtol=1e-12)
    opts = cf.SolverOptions(cfl=50.0, print_interval=0)

    class Diverging(OracleBackend):
        def run(self, U, n, scheme):
            return self.o.run(U, n, scheme, cf.step_coefficient(1, 50.0, scheme))

    U, hist, conv, iters, res, res0, msg = mr.march(U0, Diverging(), opts, 1.0)
    assert not conv and ("non-finite" in msg or "grew" in msg)


class _ScriptedBackend:
    """A backend whose residuals follow a script: the march's stop rules alone."""

    scheme = "ssprk3"

    def __init__(self, residuals, activity=float("nan")):
        self.r = list(residuals)
        self.calls = []
        self.act = activity

    def run(self, U, n, scheme):
        self.calls.append(n)
        return U, self.r.pop(0)

    def limiter_activity(self):
        return self.act


def _march(residuals, activity=float("nan"), **kw):
    cf, mr = nz_module("config"), nz_module("march")
    opts = cf.SolverOptions(print_interval=0, **kw)
    bk = _ScriptedBackend(residuals, activity)
    out = mr.march(0, bk, opts, 2.0)
    return out, bk.calls


def test_stop_rules_each_rule():
    import math

    # tolerance 1e-6 at scale 2: stops when res <= 2e-6
    (U, h, conv, it, res, res0, msg), calls = _march([1.0, 0.5, 1e-6], tolerance=1e-6,
                                                     check_interval=10)
    assert conv and it == 21 and msg == "" and calls == [1, 10, 10]
    assert h.iterations == [1, 11, 21] and res == 1e-6 and res0 == 1.0
    # converged at the first step
    (_, _, conv, it, _, _, msg), calls = _march([1e-7], tolerance=1e-6)
    assert conv and it == 1 and calls == [1]
    # non-finite first residual and later residual
    (_, _, conv, it, _, _, msg), _ = _march([math.nan])
    assert not conv and it == 1 and msg == "initial residual is not finite"
    (_, _, conv, it, _, _, msg), _ = _march([1.0, math.inf], check_interval=5)
    assert not conv and it == 6 and "non-finite after 6 iterations" in msg
    # divergence: past divergence_factor times the first residual
    (_, _, conv, it, _, _, msg), _ = _march([1.0, 2.0, 101.0], divergence_factor=100.0,
                                            check_interval=5)
    assert not conv and it == 11 and "grew by 1.0e+02x" in msg
    # budget: the last chunk is cut to the budget
    (_, _, conv, it, res, _, msg), calls = _march([1.0, 0.9, 0.5], max_iterations=15,
                                                  check_interval=10, tolerance=1e-9)
    assert not conv and it == 15 and calls == [1, 10, 4] and "max_iterations=15" in msg
    assert res == 0.5


def test_stop_rules_stall_and_precedence():
    flat = [1.0] + [0.5 + 1e-4 * k for k in range(10)]
    (_, _, conv, it, _, _, msg), _ = _march(flat, stall_window=4, stall_ratio=0.98,
                                            check_interval=3, tolerance=1e-12)
    assert not conv and it == 1 + 4 * 3 and "stalled" in msg and "smaller cfl" in msg
    (_, _, conv, _, _, _, msg), _ = _march(flat, activity=2.5, stall_window=4, stall_ratio=0.98,
                                           check_interval=3, tolerance=1e-12)
    assert "limit cycle" in msg and "~2.5 element(s)" in msg and "Refine (refine=+1)" in msg
    # a residual that meets the tolerance wins over a full stall window
    (_, _, conv, it, _, _, _), _ = _march([1.0, 1e-9, 1e-9, 1e-9], stall_window=2,
                                          tolerance=1e-9, check_interval=2)
    assert conv and it == 3
    # the window holds only the last stall_window checks
    falling = [1.0, 0.9, 0.5, 0.499, 0.498, 0.497]
    (_, _, conv, it, _, _, msg), _ = _march(falling, stall_window=3, stall_ratio=0.98,
                                            check_interval=1, tolerance=1e-12)
    assert "stalled" in msg and it == 5
