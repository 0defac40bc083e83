"""Adjoint design gradients of thrust w.r.t. the eight interior control
heights of the degree-9 Bezier wall, against central differences of the
converged thrust: tri P2 on the 14 x 7 base grid at refine 1 (every
component; forward tolerance 1e-12) and at the benchmarked refine 5 (200,704
elements: the VJP dot-product test for every flux, and the two largest
components at the reference's 1e-4 bar; forward tolerance 1e-11).

The unperturbed heights are the exact degree elevation of the smooth contour
(closed form, checked on the host); the heights here are perturbed by up to 5 %
(seed 0).  Finite differences are warm-started from the converged state.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import nz_module

pytestmark = pytest.mark.gpu


def test_bezier9_thrust_gradient_matches_central_differences():
    sens, wl, cf, ini = (nz_module("sensitivity"), nz_module("workloads"), nz_module("config"),
                         nz_module("initial"))
    geom = wl.bezier9_geometry(0)
    case = sens.DifferentiableCase(geometry=geom, order=2, refine=1, kind="tri", nx=14, nr=7,
                                   options=cf.SolverOptions(tolerance=1e-12, print_interval=0))
    U0 = ini.quasi1d_initial(case.operators(), case.flow, geom)
    names = [f"bz{i}" for i in range(1, 9)]
    J, grad, fwd, info = case.value_and_gradient("thrust", names=names, U0=U0,
                                                 adjoint_tolerance=1e-10)
    print(f"thrust {J:.10f}, forward {fwd.iterations} it, adjoint "
          f"{info['adjoint_iterations']} it", flush=True)
    base = case.default_params()
    scale = max(abs(v) for v in grad.values())
    for k in names:
        h = 1e-5
        vals = []
        for sgn in (1.0, -1.0):
            f = case.solve_forward({k: base[k] + sgn * h}, U0=fwd.U)
            assert f.converged, f.message
            vals.append(sens.objective_and_partials(f.U, f.operators, case.flow, "thrust")[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"  d thrust / d {k}: adjoint {grad[k]: .6e}  central difference {fd: .6e}",
              flush=True)
        assert grad[k] == pytest.approx(fd, rel=2e-3, abs=2e-4 * scale)


# ------------------------------------------------------------ full size
@pytest.fixture(scope="module")
def bezier9_refine5():
    """The design-gradient workload at its benchmarked size: tri P2 on the
    14 x 7 base grid at refine 5 (200,704 elements), forward tolerance 1e-11
    (this mesh's round-off plateau is about 4e-12)."""
    sens, wl, cf, ini = (nz_module("sensitivity"), nz_module("workloads"), nz_module("config"),
                         nz_module("initial"))
    geom = wl.bezier9_geometry(0)
    case = sens.DifferentiableCase(geometry=geom, order=2, refine=5, kind="tri", nx=14, nr=7,
                                   options=cf.SolverOptions(tolerance=1e-11, print_interval=0))
    ops = case.operators()
    assert ops.n_elem == 200704
    U0 = ini.quasi1d_initial(ops, case.flow, geom)
    return case, ops, U0


@pytest.mark.parametrize("flux", ["roe", "hllc", "ausmup"])
def test_vjp_dot_product_full_size(bezier9_refine5, flux):
    """<lam, (dR/dU) v> by central differences of the device residual against
    <(dR/dU)^T lam, v> from the VJP kernel, on 200,704 P2 triangles."""
    case, ops, U0 = bezier9_refine5
    cf, be = nz_module("config"), nz_module("backend")
    bk = be.HipBackend(ops, case.flow.replace(flux=flux), cf.SolverOptions())
    U, _ = bk.run(U0, 20, "ssprk3")
    rng = np.random.default_rng(3)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(0, 1)) * 1e-2
    lam = rng.standard_normal(U.shape)
    h = 1e-6
    Rp, _ = bk.residual(U + h * v)
    Rm, _ = bk.residual(U - h * v)
    lhs = float((lam * (Rp - Rm)).sum() / (2 * h))
    rhs = float((bk.residual_vjp(U, lam) * v).sum())
    print(f"{flux}: <lam, J v> {lhs:.12e}  <J^T lam, v> {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs)


def test_adjoint_gradients_full_size_against_central_differences(bezier9_refine5):
    """All eight control-height derivatives from one adjoint; the two largest
    against central differences of the converged thrust, warm-started from
    the converged state.

    With the reference test's step (h = 1e-4 relative,
    tests/test_sensitivity.py:52, src/sensitivity.py:548-549) the bar is the
    reference's 1e-4.  With h = 1e-6 relative the difference quotient
    carries the forward solves' stopping error divided by h: the solves stop
    at a scaled residual of 1e-11 (this mesh's round-off plateau is about
    4e-12), which moves thrust by about 1e-11 and the quotient by about 1e-5
    absolute, 1e-4 of these components; that case is held to 5e-4."""
    sens = nz_module("sensitivity")
    case, ops, U0 = bezier9_refine5
    names = [f"bz{i}" for i in range(1, 9)]
    J, grad, fwd, info = case.value_and_gradient("thrust", names=names, U0=U0,
                                                 adjoint_tolerance=1e-8)
    print(f"thrust {J:.12f}, forward {fwd.iterations} it, adjoint "
          f"{info['adjoint_iterations']} it; gradient {grad}", flush=True)
    assert fwd.converged and all(np.isfinite(list(grad.values())))
    base = case.default_params()
    for k in sorted(names, key=lambda n: -abs(grad[n]))[:2]:
        for step, bar in ((1e-4, 1e-4), (1e-6, 5e-4)):
            h = step * max(abs(base[k]), 1e-3)
            vals = []
            for sgn in (1.0, -1.0):
                f = case.solve_forward({k: base[k] + sgn * h}, U0=fwd.U)
                assert f.converged, f.message
                vals.append(sens.objective_and_partials(f.U, f.operators, case.flow, "thrust")[0])
            fd = (vals[0] - vals[1]) / (2 * h)
            rel = abs(grad[k] - fd) / max(abs(grad[k]), abs(fd))
            print(f"  d thrust / d {k}: adjoint {grad[k]: .9e}  central difference "
                  f"(h = {step:g} rel) {fd: .9e}  rel {rel:.2e}", flush=True)
            assert rel < bar, (k, step, rel)
