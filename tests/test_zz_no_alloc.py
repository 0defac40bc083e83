"""``nz_rate_noalloc``: the rate with its boundary pre-pass in the library's
static device scratch instead of a caller's workspace.  Through the C ABI,
against ``nz_rate`` bit for bit, on two calls in a row (the second on a
smaller mesh, so it must not depend on what the first left in the array),
then once more on the first mesh; and the size limit is enforced before any
launch."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import build_case, golden, nz_module

pytestmark = pytest.mark.gpu


def _rate_both(bk, U):
    import torch

    u = torch.from_numpy(np.ascontiguousarray(U)).to(bk.device)
    F1, w1 = torch.empty_like(u), torch.empty(bk.E, dtype=torch.float64, device=bk.device)
    F2, w2 = torch.full_like(u, np.nan), torch.full((bk.E,), np.nan, dtype=torch.float64,
                                                    device=bk.device)
    st = bk._stream()
    rc = bk.lib.nz_rate(bk._pp, bk._p(u), bk._p(F1), bk._p(w1), st)
    assert rc == 0, bk.lib.nz_last_error()
    rc = bk.lib.nz_rate_noalloc(bk._pp, bk._p(u), bk._p(F2), bk._p(w2), st)
    assert rc == 0, bk.lib.nz_last_error()
    torch.cuda.synchronize()
    return F1.cpu().numpy(), w1.cpu().numpy(), F2.cpu().numpy(), w2.cpu().numpy()


def test_noalloc_rate_matches_rate_on_consecutive_calls():
    cf, be, wl = nz_module("config"), nz_module("backend"), nz_module("workloads")
    big_ops, big_U = wl.channel_case(256, 128, 3)
    small_ops, _ = build_case("cfg1", dense=False)
    small_U = golden("cfg1")["U0"]
    cases = [(big_ops, big_U, "hllc"), (small_ops, small_U, "roe"), (big_ops, big_U, "ausmup")]
    for ops, U, flux in cases:
        bk = be.HipBackend(ops, cf.FlowConditions(flux=flux), cf.SolverOptions())
        nbq = ops.n_boundary * ops.ref.n_qface
        assert nbq <= bk.lib.nz_rate_noalloc_max_boundary_points()
        F1, w1, F2, w2 = _rate_both(bk, U)
        print(f"{ops.n_elem} elements, {nbq} boundary points, {flux}: identical "
              f"{np.array_equal(F1, F2) and np.array_equal(w1, w2)}", flush=True)
        assert np.array_equal(F1, F2) and np.array_equal(w1, w2)


def test_noalloc_rate_rejects_oversized_boundary_before_launch():
    cf, be = nz_module("config"), nz_module("backend")
    ops, _ = build_case("cfg1", dense=False)
    bk = be.HipBackend(ops, cf.FlowConditions(), cf.SolverOptions())
    limit = int(bk.lib.nz_rate_noalloc_max_boundary_points())
    q = type(bk.P).from_buffer_copy(bk.P)
    # a descriptor claiming one boundary face point more than the limit; the
    # library must refuse it on the host, so the pointers are never used
    q.n_edge = q.n_int + limit // q.nqf + 1
    dummy = C.c_void_p(1)
    rc = bk.lib.nz_rate_noalloc(C.byref(q), dummy, dummy, dummy, bk._stream())
    assert rc != 0
    assert b"static scratch" in bk.lib.nz_last_error()
