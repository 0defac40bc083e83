"""Device parity: the HIP path against the reference's golden vectors and the
CPU oracle, through the C ABI (via ``HipBackend``).

Bars: one residual evaluation <= 1e-12 relative in max norm (per flux);
rates and steps <= 1e-12; limiter activation flags bit-exact with limited
states <= 1e-13; wave-speed sums and time steps <= 1e-13.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import build_case, golden, nz_module, rel_err

pytestmark = pytest.mark.gpu

FLUXES = ("roe", "hllc", "ausmup")


def backend(ops, flow, **opt):
    cf = nz_module("config")
    be = nz_module("backend")
    return be.HipBackend(ops, flow, cf.SolverOptions(**opt))


@pytest.mark.parametrize("flux", FLUXES)
def test_residual_cfg1(flux):
    g = golden("cfg1")
    ops, flow = build_case("cfg1", dense=False)
    bk = backend(ops, flow.replace(flux=flux))
    for st in ("U0", "Ur"):
        R, w = bk.residual(g[st])
        print(flux, st, rel_err(R, g[f"R_{flux}_{st}"]), rel_err(w, g[f"wave_{flux}_{st}"]))
        assert rel_err(R, g[f"R_{flux}_{st}"]) <= 1e-12
        assert rel_err(w, g[f"wave_{flux}_{st}"]) <= 1e-13


@pytest.mark.parametrize("name,key", [("quad_q2_p2", "U0"), ("tri_p2_q2_backflow", "U"),
                                      ("synth_p3", "U"), ("synth_p1", "U"), ("tri_p0", "U0")])
@pytest.mark.parametrize("flux", FLUXES)
def test_residual_other_cases(name, key, flux):
    """Against the reference's golden residual where the fixture holds one for
    this flux; the P0 fixture holds Roe only, so P0 with HLLC and AUSM+-up is
    checked against the oracle, whose fluxes are pinned by the P1-P3 goldens
    and whose P0 path is pinned by the Roe P0 golden."""
    from oracle.oracle import CpuOracle

    g = golden(name)
    suffix = f"_{key}" if name in ("quad_q2_p2", "tri_p0") else ""
    has_golden = f"R_{flux}{suffix}" in g
    ops, flow = build_case(name, dense=not has_golden)
    flow = flow.replace(flux=flux)
    R, w = backend(ops, flow).residual(g[key])
    if has_golden:
        Rr, wr = g[f"R_{flux}{suffix}"], g[f"wave_{flux}{suffix}"]
    else:
        Rr, wr = CpuOracle(ops, flow).residual(g[key])
    print(name, flux, "golden" if has_golden else "oracle", rel_err(R, Rr))
    assert rel_err(R, Rr) <= 1e-12
    assert rel_err(w, wr) <= 1e-13


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2", "synth_p3"])
def test_rate_and_inverse_mass(name):
    from oracle.oracle import CpuOracle

    g = golden(name)
    ops, flow = build_case(name)
    U = g["U0"] if "U0" in g else g["U"]
    bk = backend(ops, flow)
    F, w = bk.rate(U)
    ref_key = {"cfg1": "rate_roe_U0", "quad_q2_p2": "rate_hllc_U0", "synth_p3": "rate_roe"}[name]
    assert rel_err(F, g[ref_key]) <= 1e-12
    orc = CpuOracle(ops, flow)
    Fo, wo = orc.rate(U)
    assert rel_err(F, Fo) <= 1e-12 and rel_err(w, wo) <= 1e-13
    R, _ = bk.residual(U)
    assert rel_err(-bk.inverse_mass(R), g[ref_key]) <= 1e-12


def test_time_step_and_norm():
    g = golden("cfg1")
    ops, flow = build_case("cfg1", dense=False)
    bk = backend(ops, flow)
    F, w = bk.rate(g["U0"])
    assert rel_err(bk.time_step(w)[:, 0, 0], g["dt_U0"]) <= 1e-13
    assert abs(bk.norm(F) / float(g["norm_rate_roe_U0"]) - 1) <= 1e-13


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2"])
def test_limiter_flags_bit_exact(name):
    from oracle.oracle import CpuOracle

    g = golden(name)
    ops, flow = build_case(name)
    bk = backend(ops, flow)
    out, flags, (ns, nr) = bk.limit_flags(g["Us"])
    assert np.array_equal(flags, g["Us_flags"])
    assert rel_err(out, g["Us_positivity"]) <= 1e-13
    o2, f2, c2 = CpuOracle(ops, flow).limit(g["Us"])
    assert np.array_equal(flags, f2) and (ns, nr) == c2
    assert rel_err(out, o2) <= 1e-15
    # inactive on an admissible state: bit-identical output
    out, flags, cnt = bk.limit_flags(g["U0"])
    assert np.array_equal(out, g["U0"]) and not flags.any() and cnt == (0, 0)


@pytest.mark.parametrize("scheme,steps,key", [("ssprk3", 1, "ssprk3_1step"),
                                              ("ssprk3", 10, "ssprk3_10step"),
                                              ("euler", 1, "euler_1step")])
def test_steps_cfg1(scheme, steps, key):
    g = golden("cfg1")
    ops, flow = build_case("cfg1", dense=False)
    bk = backend(ops, flow, scheme=scheme)
    U, res = bk.run(g["U0"], steps, scheme)
    print(scheme, steps, rel_err(U, g[f"{key}_U"]), res, float(g[f"{key}_res"]))
    assert rel_err(U, g[f"{key}_U"]) <= 1e-12
    assert abs(res / float(g[f"{key}_res"]) - 1) <= 1e-12


@pytest.mark.parametrize("flux", FLUXES)
def test_steps_synth_p3_against_oracle(flux):
    """The fused stage kernel at P3 on affine triangles (the instantiation
    that marches the 1024x512 channel): two SSP-RK3 steps against the CPU
    oracle, limiter on, with the limiter's counts."""
    from oracle.oracle import CpuOracle

    g = golden("synth_p3")
    ops, flow = build_case("synth_p3")
    flow = flow.replace(flux=flux)
    bk = backend(ops, flow)
    U, res = bk.run(g["U"], 2, "ssprk3")
    Uo, reso = CpuOracle(ops, flow).run(g["U"], 2, "ssprk3", bk.dt_factor("ssprk3"))
    print(flux, rel_err(U, Uo), res, reso, bk.limiter_activity())
    assert rel_err(U, Uo) <= 1e-12
    assert abs(res / reso - 1) <= 1e-12


def test_steps_quad_p2_and_p0():
    g = golden("quad_q2_p2")
    ops, flow = build_case("quad_q2_p2", dense=False)
    U, res = backend(ops, flow).run(g["U0"], 1, "ssprk3")
    assert rel_err(U, g["ssprk3_1step_U"]) <= 1e-12
    assert abs(res / float(g["ssprk3_1step_res"]) - 1) <= 1e-12
    g = golden("tri_p0")
    ops, flow = build_case("tri_p0", dense=False)
    bk = backend(ops, flow)
    assert bk.scheme == "euler"
    U, res = bk.run(g["U0"], 5, bk.scheme)
    assert rel_err(U, g["5step_U"]) <= 1e-12
    assert abs(res / float(g["5step_res"]) - 1) <= 1e-12


def test_run_chunking_and_no_alias():
    """Chunking invariance (<=1e-13) and the result never aliases backend buffers."""
    g = golden("cfg1")
    ops, flow = build_case("cfg1", dense=False)
    bk = backend(ops, flow)
    a, ra = bk.run(g["U0"], 6, "ssprk3")
    b, _ = bk.run(g["U0"], 2, "ssprk3")
    b, rb = bk.run(b, 4, "ssprk3")
    assert rel_err(a, b) <= 1e-13 and abs(ra - rb) <= 1e-13 * abs(ra)
    import torch

    t = bk.asarray(g["U0"])
    out, _ = bk.run(t, 2, "ssprk3")
    for buf in bk._S:
        assert out.data_ptr() != buf.data_ptr()


def test_unknown_inputs_fail_loudly():
    cf = nz_module("config")
    be = nz_module("backend")
    ops, flow = build_case("cfg1", dense=False)
    bk = be.HipBackend(ops, flow, cf.SolverOptions())
    with pytest.raises(ValueError, match="unknown scheme"):
        bk.run(golden("cfg1")["U0"], 1, "rk4")
    with pytest.raises(ValueError, match="unknown scheme"):
        bk.make_step("rk4")


@pytest.mark.parametrize("tvb,key", [(0.0, "Ub_superbee"), (50.0, "Ub_superbee_tvb50")])
def test_superbee_alone(tvb, key):
    import ctypes as C

    import torch

    g = golden("cfg1")
    ops, flow = build_case("cfg1", dense=False)
    bk = backend(ops, flow, limiter="superbee", tvb_constant=tvb)
    u = bk.asarray(g["Ub"]).clone()
    rc = bk.lib.nz_superbee(bk._pp, C.c_void_p(u.data_ptr()), tvb, bk.area_total,
                            C.c_void_p(bk._sb_work.data_ptr()), None, bk._stream())
    assert rc == 0
    torch.cuda.synchronize()
    assert rel_err(u.cpu().numpy(), g[key]) <= 1e-13


def test_superbee_then_positivity_and_steps():
    from oracle.oracle import CpuOracle

    g = golden("cfg1")
    ops, flow = build_case("cfg1")
    bk = backend(ops, flow, limiter="superbee", tvb_constant=0.0)
    assert rel_err(bk.limit(g["Ub"]), g["Ub_apply_superbee"]) <= 1e-13
    U, res = bk.run(g["Ub"], 3, "ssprk3")
    Uo, reso = CpuOracle(ops, flow).run(g["Ub"], 3, "ssprk3", bk.dt_factor(), "superbee", 0.0)
    assert rel_err(U, Uo) <= 1e-12 and abs(res / reso - 1) <= 1e-12


@pytest.mark.parametrize("q", [1, 2])
@pytest.mark.parametrize("flux", FLUXES)
def test_quad_p3_against_oracle(q, flux):
    """Quadrilaterals at P3 (N = 16), straight (Q1) and curved (Q2) edges, on
    the quasi-1D initial state of the smooth nozzle at refine 1 (280
    elements): the rate, then two SSP-RK3 steps with the positivity limiter,
    against the CPU oracle."""
    from oracle.oracle import CpuOracle

    mesh, op, cf, ini = (nz_module(m) for m in ("mesh", "operators", "config", "initial"))
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind="quad", refine=1, geometry_order=q)
    ops = op.build_operators(topo, coords, 3, dense=True)
    flow = cf.FlowConditions().replace(flux=flux)
    U0 = ini.quasi1d_initial(ops, flow, geom)
    bk = backend(ops, flow)
    orc = CpuOracle(ops, flow)
    F, w = bk.rate(U0)
    Fo, wo = orc.rate(U0)
    print("quad P3 Q", q, flux, rel_err(F, Fo), rel_err(w, wo))
    assert rel_err(F, Fo) <= 1e-12 and rel_err(w, wo) <= 1e-13
    U, res = bk.run(U0, 2, "ssprk3")
    Uo, reso = orc.run(U0, 2, "ssprk3", bk.dt_factor("ssprk3"))
    print("  two steps", rel_err(U, Uo), res, reso, bk.limiter_activity())
    assert rel_err(U, Uo) <= 1e-12
    assert abs(res / reso - 1) <= 1e-12


@pytest.mark.parametrize("name", ["cfg1", "quad_q2_p2", "tri_p0"])
def test_device_projection_forms_the_reference_initial_state(name):
    """``nz_project`` (``initialize.py:41-46``) forms the reference's own
    quasi-1D initial state (fixture ``U0``) to 1e-13, and projects random
    point values as the host projection does."""
    from conftest import CASES

    ini, mesh = nz_module("initial"), nz_module("mesh")
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    mk = CASES[name][0]
    geom = mesh.NozzleGeometry(contour=mk["contour"], area_ratio=mk["area_ratio"])
    bk = backend(ops, flow)
    U0 = ini.quasi1d_initial(ops, flow, geom, backend=bk)
    assert isinstance(U0, np.ndarray) and U0.shape == g["U0"].shape
    print(name, rel_err(U0, g["U0"]))
    assert rel_err(U0, g["U0"]) < 1e-13
    vals = np.random.default_rng(2).standard_normal((ops.n_elem, ops.ref.n_qvol, 4))
    assert rel_err(bk.project(vals), ini.project(vals, ops)) < 1e-13


@pytest.mark.parametrize("kind,q,p", [("quad", 1, 3), ("quad", 2, 3), ("tri", 2, 2), ("tri", 1, 3)])
def test_device_projection_is_exact_on_the_space(kind, q, p):
    """Point values of a random element of the space project back onto it."""
    mesh, op, cf = nz_module("mesh"), nz_module("operators"), nz_module("config")
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind=kind, refine=1, geometry_order=q)
    ops = op.build_operators(topo, coords, p, dense=False)
    bk = backend(ops, cf.FlowConditions())
    U = np.random.default_rng(4).standard_normal((ops.n_elem, ops.ref.n_basis, 4))
    vals = np.einsum("iq,eis->eqs", np.asarray(ops.ref.phi_vol), U)
    assert rel_err(bk.project(vals), U) < 1e-12
    with pytest.raises(ValueError, match="shape"):
        bk.project(vals[:, :-1])


def test_device_projection_uniform_start_and_order_change():
    """The uniform start and the order change (``initialize.py:61-65,
    121-147``) through the device projection: the same states as on the host
    (1e-13), and a P1 field raised to P2 and lowered again comes back."""
    ini, mesh, op, cf = (nz_module(m) for m in ("initial", "mesh", "operators", "config"))
    ops, flow = build_case("cfg1", dense=False)
    bk = backend(ops, flow)
    assert rel_err(ini.uniform_initial(ops, flow, backend=bk), ini.uniform_initial(ops, flow)) < 1e-13
    geom = mesh.NozzleGeometry(contour="smooth", area_ratio=2.5)
    topo, coords, _ = mesh.build_nozzle_mesh(geom, kind="quad", refine=1, geometry_order=2)
    o1 = op.build_operators(topo, coords, 1, dense=False)
    o2 = op.build_operators(topo, coords, 2, dense=False)
    b1, b2 = backend(o1, cf.FlowConditions()), backend(o2, cf.FlowConditions())
    U1 = np.random.default_rng(6).standard_normal((o1.n_elem, o1.ref.n_basis, 4))
    U2 = ini.change_order(U1, o1, o2, backend=b2)
    assert rel_err(U2, ini.change_order(U1, o1, o2)) < 1e-13
    assert rel_err(ini.change_order(U2, o2, o1, backend=b1), U1) < 1e-12


def test_device_observables_of_the_reference_state():
    """``nz_boundary_integrals`` on the reference's converged state gives the
    reference's thrust, wall thrust, mass flows, exit Mach and exit pressure
    (``src/post/postprocess.py:193-226``) to 1e-13, as the host observables
    do; and on a curved quadrilateral state the two agree."""
    ob = nz_module("observables")
    g = golden("cfg1_converged")
    ops, flow = build_case("cfg1", dense=False)
    perf = backend(ops, flow).performance(g["U"])
    for k in ("thrust", "thrust_wall", "mass_flow_in", "mass_flow_out",
              "exit_mach_area_averaged", "exit_pressure_ratio"):
        print(k, getattr(perf, k), float(g[f"perf_{k}"]))
        assert abs(getattr(perf, k) / float(g[f"perf_{k}"]) - 1) <= 1e-13, k
    gq = golden("quad_q2_p2")
    ops, flow = build_case("quad_q2_p2", dense=False)
    a = backend(ops, flow).performance(gq["U0"])
    b = ob.performance(gq["U0"], ops, flow)
    for k in ("thrust", "thrust_wall", "mass_flow_in", "mass_flow_out",
              "exit_mach_area_averaged", "exit_pressure_ratio"):
        assert abs(getattr(a, k) - getattr(b, k)) <= 1e-13 * abs(getattr(b, k)), k
