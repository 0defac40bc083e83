"""The shocked quadrilateral march (quads, Q2, P2, bell contour, pb 0.7, HLLC):
limiter activation flags on the device against the CPU oracle, bit for bit,
on states of a real march.

The case is the reference's shocked setup at refine 3 (4,480 elements) and
at the benchmarked refine 7 (1,146,880 elements).  After 20 SSP-RK3 steps
from the quasi-1D initial state (which holds the normal shock), one stage is
taken without the limiter; the same input then goes to the fused stage
limiter, the standalone device limiter and the oracle's limiter.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import nz_module, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[3, 7], ids=["refine3", "refine7"])
def shocked(request):
    import torch

    mesh, op, cf = nz_module("mesh"), nz_module("operators"), nz_module("config")
    ini, be = nz_module("initial"), nz_module("backend")
    geom = mesh.NozzleGeometry(contour="bell", area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind="quad", refine=request.param,
                                             geometry_order=2)
    ops = op.build_operators(topo, coords, 2, dense=False)
    flow = cf.FlowConditions(back_pressure_ratio=0.7, allow_shock_in_nozzle=True, flux="hllc")
    lim = be.HipBackend(ops, flow, cf.SolverOptions(limiter="positivity"))
    nolim = be.HipBackend(ops, flow, cf.SolverOptions(limiter="none"))
    U0 = ini.quasi1d_initial(ops, flow, geom, backend=lim)
    U, _ = lim.run(U0, 20, "ssprk3")
    print(f"shocked quad Q2 P2: {ops.n_elem} elements, limited so far {lim.n_limited}", flush=True)
    return ops, flow, lim, nolim, torch.from_numpy(np.asarray(U)).to(lim.device)


def _stage(bk, u, out, flags):
    bk._chk(bk.lib.nz_stage(bk._pp, bk._p(u), None, bk._p(out), bk._p(bk._dt),
                            bk.dt_factor("ssprk3"), 1.0, 1.0, 1, bk._p(bk._scratch),
                            C.c_void_p(flags.data_ptr()) if flags is not None else None,
                            bk._p(bk._counters), bk._stream()), "nz_stage")


def test_limiter_flags_match_oracle_on_a_shocked_march(shocked):
    import torch

    from oracle.oracle import positivity_limit

    ops, flow, lim, nolim, u = shocked
    ustar = torch.empty_like(u)
    _stage(nolim, u, ustar, None)  # the stage's state before limiting
    ulim = torch.empty_like(u)
    flags = torch.zeros(ops.n_elem, dtype=torch.uint8, device=u.device)
    lim._counters.zero_()
    _stage(lim, u, ulim, flags)  # the fused stage limiter on the same input
    torch.cuda.synchronize()
    counts = lim._counters.cpu().tolist()
    Ustar = ustar.cpu().numpy()
    ref_U, ref_flags, (n_scaled, n_rep) = positivity_limit(Ustar, ops, flow)
    g_flags = flags.cpu().numpy()
    print(f"flags set: oracle {int((ref_flags != 0).sum())}, device {int((g_flags != 0).sum())}, "
          f"scaled {n_scaled}, repaired {n_rep}", flush=True)
    assert n_scaled > 0  # the shock keeps the limiter busy
    assert np.array_equal(g_flags, ref_flags)
    # the activity counters (summed per wave on the device) count the same elements
    assert (int(counts[0]), int(counts[1])) == (int(n_scaled), int(n_rep)), counts
    assert rel_err(ulim.cpu().numpy(), ref_U) < 1e-13
    # the standalone device limiter on the same input
    lu, lflags, _ = lim.limit_flags(Ustar)
    assert np.array_equal(np.asarray(lflags), ref_flags)
    assert rel_err(np.asarray(lu), ref_U) < 1e-13
