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
