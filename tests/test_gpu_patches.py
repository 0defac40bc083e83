"""The element kernel on patch tiles (``nz_problem.patches``): 16 elements
that share edges per workgroup, each shared edge's Riemann flux solved once.

Bars: the residual, the rate, the wave-speed sums and every stage state are
bit-identical to the tiled kernel's (the patch kernel forms each trace and
flux from the same operands in the same order, and the flux of an interior
edge is the same number from either side); the residual norm, whose partial
sums are grouped by workgroup, agrees to 1e-14 relative; the residual matches
the CPU oracle to 1e-12 relative.  Patches of randomly chosen elements (every
face off the patch, more than the 24 ghost rows) exercise the overflow path.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import nz_module, rel_err

pytestmark = pytest.mark.gpu

FLUXES = ("roe", "hllc", "ausmup")


def _case(order, flux="roe", ncx=128, ncy=64, limiter="positivity", dense=False):
    # dense operators: the oracle's form
    cf, be, wl = nz_module("config"), nz_module("backend"), nz_module("workloads")
    ops, U = wl.channel_case(ncx, ncy, order, dense=dense)
    flow = cf.FlowConditions(flux=flux)
    bk = be.HipBackend(ops, flow, cf.SolverOptions(limiter=limiter))
    return ops, flow, bk, U


def _random_patches(bk, seed=0):
    be = nz_module("backend")
    order = np.random.default_rng(seed).permutation(bk.E)
    bk._patches = bk._torch.from_numpy(be.patch_table(order).reshape(-1)).to(bk.device)
    bk.set_patches(True)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("flux", FLUXES)
def test_patch_residual_and_rate_bit_identical(order, flux):
    ops, flow, bk, U = _case(order, flux, dense=order in (1, 3) and flux == "roe")
    bk.set_patches(True)
    Rp, wp = bk.residual(U)
    Fp, _ = bk.rate(U)
    bk.set_patches(False)
    Rt, wt = bk.residual(U)
    Ft, _ = bk.rate(U)
    print(order, flux, "max |dR|", float(np.abs(Rp - Rt).max()), "max |dF|",
          float(np.abs(Fp - Ft).max()))
    assert np.array_equal(Rp, Rt) and np.array_equal(wp, wt) and np.array_equal(Fp, Ft)
    if order in (1, 3) and flux == "roe":
        from oracle.oracle import CpuOracle

        Ro, wo = CpuOracle(ops, flow).residual(U)
        assert rel_err(Rp, Ro) <= 1e-12 and rel_err(wp, wo) <= 1e-13


@pytest.mark.parametrize("order", [1, 3])
def test_patch_march_bit_identical(order):
    """Ten SSP-RK3 steps (fused stages, positivity limiter) on patches and on
    tiles: the same states; the residual norm to its summation order."""
    _, _, bk, U = _case(order)
    bk.set_patches(True)
    Up, rp = bk.run(U, 10, "ssprk3")
    bk.set_patches(False)
    Ut, rt = bk.run(U, 10, "ssprk3")
    print(order, "residual", rp, rt)
    assert np.array_equal(Up, Ut)
    assert abs(rp - rt) <= 1e-14 * abs(rt)


@pytest.mark.parametrize("order", [0, 3])
def test_random_patches_overflow_the_ghost_rows(order):
    """Patches of 16 random elements: nearly every face is off the patch, so
    most off-patch faces form the neighbour's trace in the flux item."""
    _, _, bk, U = _case(order, ncx=96, ncy=96)
    bk.set_patches(False)
    Rt, wt = bk.residual(U)
    Ut, _ = bk.run(U, 3, "ssprk3")
    _random_patches(bk)
    Rp, wp = bk.residual(U)
    Up, _ = bk.run(U, 3, "ssprk3")
    assert np.array_equal(Rp, Rt) and np.array_equal(wp, wt)
    assert np.array_equal(Up, Ut)



def test_element_kernel_autotuner():
    """The autotuner times both work shapes on the mesh, keeps one per (kind,
    order, flux, structured) and every later backend of that key uses it."""
    be = nz_module("backend")
    _, _, bk, U = _case(1)
    choice = bk.tune_element_kernel(U)
    assert isinstance(choice, bool) and bk.patches == choice
    _, _, bk2, _ = _case(1)
    assert bk2.patches == choice
    assert be._DG_PATCHES[bk._dg_key] == choice
    _, _, small, Us = _case(1, ncx=32, ncy=16)  # below the locality threshold: tiles
    assert small.tune_element_kernel(Us) is False and not small.patches
