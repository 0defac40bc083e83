"""Adjoint design gradients of thrust w.r.t. the eight interior control
heights of the degree-9 Bezier wall, against central differences of the
converged thrust (tri P2, 14 x 7 base grid, refine 1; forward tolerance 1e-12).

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
