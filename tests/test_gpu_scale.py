"""Device parity at the benchmarked size of the residual microbenchmark: the
synthetic channel, P3 triangles, 256 x 128 cells (65,536 elements) and
1024 x 512 cells (2^20 elements, the benchmark's workload).

Meshes of 2^14 elements or more launch the element kernel through the
locality tile table (Morton-ordered tiles dealt to the L2 domains), which the
small golden cases never reach.  Bars: the locality-ordered launch and plain
consecutive tiling agree bit for bit (rate, wave-speed sums, two SSP-RK3
steps with the positivity limiter, and the residual norm); the device agrees
with the CPU oracle to 1e-12 relative in max norm (rates and states) and
1e-13 (wave-speed sums), with identical limiter counts.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import nz_module, rel_err

pytestmark = pytest.mark.gpu

FLUXES = ("roe", "hllc", "ausmup")


def _backend(ops, flux, **opt):
    cf, be = nz_module("config"), nz_module("backend")
    return be.HipBackend(ops, cf.FlowConditions(flux=flux), cf.SolverOptions(**opt))


def _plain(bk):
    """The same backend with consecutive tiles of 16 (no tile table)."""
    bk.P.tiles, bk.P.n_tiles = C.c_void_p(0), 0
    bk._graphs.clear()
    return bk


@pytest.fixture(scope="module")
def channel_65k():
    wl = nz_module("workloads")
    ops, U = wl.channel_case(256, 128, 3, dense=True)
    assert ops.n_elem == 65536
    return ops, U


@pytest.fixture(scope="module")
def channel_full():
    wl = nz_module("workloads")
    ops, U = wl.channel_case(1024, 512, 3, dense=True)
    assert ops.n_elem == 1 << 20
    print(f"channel 1024x512 P3: {ops.n_elem} elements, {ops.n_edges_total} edges", flush=True)
    return ops, U


@pytest.mark.parametrize("flux", FLUXES)
def test_channel_65k_rate_and_steps_against_oracle(channel_65k, flux):
    from oracle.oracle import CpuOracle

    ops, U = channel_65k
    bk = _backend(ops, flux)
    assert bk._tiles is not None  # the locality tile table is in use
    orc = CpuOracle(ops, bk.flow)
    F, w = bk.rate(U)
    Fo, wo = orc.rate(U)
    print(flux, "rate", rel_err(F, Fo), "wave", rel_err(w, wo), flush=True)
    assert rel_err(F, Fo) <= 1e-12 and rel_err(w, wo) <= 1e-13
    U2, res = bk.run(U, 2, "ssprk3")
    U2o, reso = orc.run(U, 2, "ssprk3", bk.dt_factor())
    print(flux, "2 steps", rel_err(U2, U2o), res, reso, "limited", bk.n_limited, flush=True)
    assert rel_err(U2, U2o) <= 1e-12
    assert abs(res / reso - 1) <= 1e-12


@pytest.mark.parametrize("flux", FLUXES)
def test_full_size_locality_tiles_bit_identical_and_oracle(channel_full, flux):
    """2^20 elements: the benchmark's launch (locality tile table) against
    plain tiling bit for bit, and against the oracle."""
    from oracle.oracle import CpuOracle

    ops, U = channel_full
    loc = _backend(ops, flux)
    plain = _plain(_backend(ops, flux))
    assert loc._tiles is not None
    F1, w1 = loc.rate(U)
    F2, w2 = plain.rate(U)
    assert np.array_equal(F1, F2) and np.array_equal(w1, w2)
    Fo, wo = CpuOracle(ops, loc.flow).rate(U)
    print(flux, "full-size rate vs oracle", rel_err(F1, Fo), rel_err(w1, wo), flush=True)
    assert rel_err(F1, Fo) <= 1e-12 and rel_err(w1, wo) <= 1e-13
    R1, _ = loc.residual(U)
    R2, _ = plain.residual(U)
    assert np.array_equal(R1, R2)


def test_full_size_steps_bit_identical_and_oracle(channel_full):
    """Two fused SSP-RK3 steps (positivity limiter on) at 2^20 elements: the
    locality-ordered launch against plain tiling bit for bit (state, norm,
    limiter counts), and against the oracle."""
    from oracle.oracle import CpuOracle

    ops, U = channel_full
    loc = _backend(ops, "roe")
    plain = _plain(_backend(ops, "roe"))
    U1, r1 = loc.run(U, 2, "ssprk3")
    U2, r2 = plain.run(U, 2, "ssprk3")
    assert np.array_equal(U1, U2) and r1 == r2
    assert (loc.n_limited, loc.n_mean_repaired) == (plain.n_limited, plain.n_mean_repaired)
    Uo, ro = CpuOracle(ops, loc.flow).run(U, 2, "ssprk3", loc.dt_factor())
    print("full-size 2 steps vs oracle", rel_err(U1, Uo), r1, ro, "limited", loc.n_limited,
          flush=True)
    assert rel_err(U1, Uo) <= 1e-12 and abs(r1 / ro - 1) <= 1e-12


@pytest.mark.parametrize("flux", FLUXES)
def test_adjoint_on_locality_tiles_bit_identical(channel_65k, flux):
    """The one-wave adjoint kernel takes its tiles from the element kernel's
    locality table when that table covers the mesh (``tiles_cover_mesh``):
    the VJP, three adjoint RK4 steps and the adjoint's defect equal plain
    tiling's bit for bit (the table's slots keep the defect's partial sums in
    element order)."""
    be = nz_module("backend")
    ops, U = channel_65k
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(U.shape)
    seed = rng.standard_normal(U.shape) * 1e-3
    out = []
    for plain in (False, True):
        bk = _backend(ops, flux)
        assert bk.P.tiles_cover_mesh == 1 and bk._tiles is not None
        if plain:
            _plain(bk)
        key = bk._vjp_key
        had = be._VJP_KERNEL.get(key)
        be._VJP_KERNEL[key] = 0  # the one-wave kernel, the one that reads the table
        try:
            v = bk.residual_vjp(U, lam)
            la, rel, its, _ = bk.adjoint_solve(U, seed, tolerance=1e-30, max_iterations=3,
                                               check_interval=3)
        finally:
            if had is None:
                be._VJP_KERNEL.pop(key, None)
            else:
                be._VJP_KERNEL[key] = had
        out.append((v, la, rel, its))
    (v0, l0, r0, i0), (v1, l1, r1, i1) = out
    print(flux, "defect", r0, r1, flush=True)
    assert np.array_equal(v0, v1)
    assert np.array_equal(l0, l1)
    assert r0 == r1 and i0 == i1 == 3
