"""The residual VJP kernel and the device adjoint (MI355X).

* dot-product tests: ``<lam, dR/dU v> = <vjp(lam), v>`` with ``dR/dU v`` by
  central differences of the device residual, for every flux and boundary
  kind (FD truncation bounds the tolerance at 1e-6 relative);
* geometry and back-pressure cotangents against central differences of the
  residual under node motion and back-pressure change;
* the residual as a torch autograd operation (``sensitivity.residual``):
  backward equal to the direct VJP calls, node gradient against central
  differences;
* the captured sensitivity case of the reference (smooth contour, p=1,
  refine 0, p_b/p_t = 0.15, tolerance 1e-11, Roe): thrust 0.0458513361 and
  the adjoint derivatives printed in docs/validation/gradients.md:12-20, to
  the 7 digits printed there; and central finite differences of the
  converged objective (h = 1e-4 relative, as the reference example) for area_ratio.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import build_case, golden, nz_module

pytestmark = pytest.mark.gpu

FLUXES = ("roe", "hllc", "ausmup")


def backend(ops, flow, **opt):
    cf = nz_module("config")
    be = nz_module("backend")
    return be.HipBackend(ops, flow, cf.SolverOptions(**opt))


def _jvp_fd(bk, U, v, h=1e-6):
    Rp, _ = bk.residual(U + h * v)
    Rm, _ = bk.residual(U - h * v)
    return (Rp - Rm) / (2 * h)


@pytest.mark.parametrize("name,key", [("cfg1", "U0"), ("cfg1", "Ur"), ("quad_q2_p2", "U0"),
                                      ("tri_p2_q2_backflow", "U"), ("synth_p1", "U")])
@pytest.mark.parametrize("flux", FLUXES)
def test_vjp_dot_product(name, key, flux):
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    bk = backend(ops, flow.replace(flux=flux))
    U = g[key]
    rng = np.random.default_rng(7)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(axis=(0, 1)) * 1e-2
    lam = rng.standard_normal(U.shape)
    lhs = float((lam * _jvp_fd(bk, U, v)).sum())
    rhs = float((bk.residual_vjp(U, lam) * v).sum())
    print(name, key, flux, lhs, rhs)
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs)


@pytest.mark.parametrize("name,key", [("cfg1", "U0"), ("quad_q2_p2", "U0"),
                                      ("tri_p2_q2_backflow", "U")])
def test_geometry_and_back_pressure_cotangents(name, key):
    op = nz_module("operators")
    sens = nz_module("sensitivity")
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    U = g[key]
    bk = backend(ops, flow)
    rng = np.random.default_rng(11)
    lam = rng.standard_normal(U.shape)
    _, gmet, gedge, gpb = bk.residual_vjp(U, lam, geometry=True)
    gX = sens.metric_vjp(ops, gmet) + sens.edge_vjp(ops, gedge)
    X = np.asarray(ops.node_coords)
    dX = rng.standard_normal(X.shape) * 1e-2
    h = 1e-6

    def lamR(Xn, fl=flow):
        o = op.build_operators(ops.topology, Xn, ops.ref.order, dense=False)
        R, _ = backend(o, fl).residual(U)
        return float((lam * R).sum())

    fd = (lamR(X + h * dX) - lamR(X - h * dX)) / (2 * h)
    an = float((gX * dX).sum())
    print(name, "geometry", fd, an)
    assert abs(fd - an) <= 1e-6 * abs(an)
    pb = flow.back_pressure_ratio
    hp = 1e-6
    fdp = (lamR(X, flow.replace(back_pressure_ratio=pb + hp))
           - lamR(X, flow.replace(back_pressure_ratio=pb - hp))) / (2 * hp)
    print(name, "back pressure", fdp, gpb)
    assert abs(fdp - gpb) <= 1e-6 * max(abs(gpb), 1e-3 * np.abs(lam).max())


@pytest.mark.parametrize("name,key", [("cfg1", "U0"), ("quad_q2_p2", "U0"),
                                      ("tri_p2_q2_backflow", "U")])
def test_autograd_residual(name, key):
    """``sensitivity.residual`` as a torch autograd operation: forward is the
    device residual, backward the device VJP bit for bit, and the node
    gradient matches central differences of <g, R> under node motion."""
    import torch

    op = nz_module("operators")
    sens = nz_module("sensitivity")
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    bk = backend(ops, flow)
    rng = np.random.default_rng(13)
    U = torch.tensor(g[key], device=bk.device, requires_grad=True)
    X = torch.tensor(np.asarray(ops.node_coords), dtype=torch.float64, requires_grad=True)
    gR = torch.tensor(rng.standard_normal(g[key].shape), device=bk.device)
    R = sens.residual(bk, U, X)
    assert torch.equal(R, bk.residual(U.detach())[0])
    (R * gR).sum().backward()
    # with X the geometry launch runs (dual-number Jacobians): same kernel, same bits
    assert torch.equal(U.grad, bk.residual_vjp(U.detach(), gR, geometry=True)[0])
    assert X.grad.shape == X.shape and X.grad.device == X.device
    # state only: X left out, the state-cotangent launch (reverse-mode Roe).
    # The two launches differentiate the Roe flux in different orders; they
    # agree to 3e-16 on the regular states.  The backflow state has elements
    # whose pressure is negative at every point (floored); there the Roe-mean
    # sound speed is floored to 1e-5, the wave decomposition divides by c^2
    # and cancels, rounding differences are amplified, and the two differ by
    # 3e-9 relative.
    U2 = U.detach().clone().requires_grad_(True)
    (sens.residual(bk, U2) * gR).sum().backward()
    assert torch.equal(U2.grad, bk.residual_vjp(U.detach(), gR))
    tol = 1e-8 if "backflow" in name else 1e-15
    assert float((U2.grad - U.grad).abs().max() / U.grad.abs().max()) <= tol
    Xn = np.asarray(ops.node_coords)
    dX = rng.standard_normal(Xn.shape) * 1e-2
    h = 1e-6
    lam = gR.cpu().numpy()

    def lamR(Xp):
        o = op.build_operators(ops.topology, Xp, ops.ref.order, dense=False)
        Rp, _ = backend(o, flow).residual(g[key])
        return float((lam * Rp).sum())

    fd = (lamR(Xn + h * dX) - lamR(Xn - h * dX)) / (2 * h)
    an = float((X.grad.numpy() * dX).sum())
    print(name, "autograd geometry", fd, an)
    assert abs(fd - an) <= 1e-6 * abs(an)
    with pytest.raises(ValueError, match="differs"):
        sens.residual(bk, U.detach(), X.detach() + 1.0)
    with pytest.raises(ValueError, match="float64 tensor"):
        sens.residual(bk, U.detach().cpu())


def _case():
    sens = nz_module("sensitivity")
    mesh = nz_module("mesh")
    cf = nz_module("config")
    g = golden("adjoint_case")
    dc = sens.DifferentiableCase(geometry=mesh.NozzleGeometry(contour="smooth"),
                                 flow=cf.FlowConditions(back_pressure_ratio=0.15),
                                 order=1, refine=0,
                                 options=cf.SolverOptions(tolerance=1e-11, print_interval=0))
    assert np.array_equal(dc.coords(), g["node_coords"])
    return dc, g


def test_captured_adjoint_gradients():
    """docs/validation/gradients.md:12-20: thrust and five adjoint derivatives."""
    dc, g = _case()
    names = ("area_ratio", "throat_x", "theta_exit", "inlet_half_height", "back_pressure")
    J, grad, fwd, info = dc.value_and_gradient("thrust", names=names, U0=g["U0"])
    print("J", J, "grad", grad, info, "iterations", fwd.iterations)
    assert fwd.iterations == int(g["iterations"])
    assert abs(J - float(g["thrust"])) <= 1e-10 * abs(J)
    assert abs(J - 0.0458513361) <= 1e-10
    captured = {"area_ratio": 9.772082e-03, "throat_x": -1.490300e-02,
                "theta_exit": 7.602688e-03, "inlet_half_height": -1.524949e+00}
    for k, v in captured.items():
        assert abs(grad[k] - v) <= 6e-7 * abs(v) + 1e-12, (k, grad[k], v)
    assert abs(grad["back_pressure"]) <= 1e-12


def test_gradient_against_central_differences():
    dc, g = _case()
    J, grad, fwd, _ = dc.value_and_gradient("thrust", names=("area_ratio", "length"), U0=g["U0"])
    sens = nz_module("sensitivity")
    base = dc.default_params()
    for k in ("area_ratio", "length"):
        h = 1e-4 * abs(base[k])
        vals = []
        for s in (+1, -1):
            res = dc.solve_forward({k: base[k] + s * h}, U0=fwd.U)
            assert res.converged
            vals.append(sens.objective_and_partials(res.U, res.operators, dc.flow_at(), "thrust")[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        print(k, "adjoint", grad[k], "fd", fd)
        assert abs(fd - grad[k]) <= 1e-4 * abs(grad[k])
    # the reference's recorded length derivative (docs/validation/gradients.md:28)
    assert abs(grad["length"] - 1.353855e-2) <= 6e-7 * 1.353855e-2


def test_gmres_adjoint_matches_the_march():
    """Restarted GMRES on the same linear system: the captured derivatives to
    their printed digits, and the march's gradient to the adjoint tolerance,
    in far fewer VJPs than the march's four per step."""
    dc, g = _case()
    names = ("area_ratio", "throat_x", "theta_exit", "inlet_half_height")
    _, grad_rk, _, info_rk = dc.value_and_gradient("thrust", names=names, U0=g["U0"])
    _, grad_gm, _, info_gm = dc.value_and_gradient("thrust", names=names, U0=g["U0"],
                                                   adjoint_method="gmres")
    print("rk4", info_rk, "gmres", info_gm)
    assert info_gm["adjoint_relative_residual"] <= 1e-8
    assert info_gm["adjoint_iterations"] < 4 * info_rk["adjoint_iterations"]
    captured = {"area_ratio": 9.772082e-03, "throat_x": -1.490300e-02,
                "theta_exit": 7.602688e-03, "inlet_half_height": -1.524949e+00}
    for k, v in captured.items():
        assert abs(grad_gm[k] - v) <= 6e-7 * abs(v) + 1e-12, (k, grad_gm[k], v)
        assert abs(grad_gm[k] - grad_rk[k]) <= 1e-6 * abs(grad_rk[k]), (k, grad_gm[k], grad_rk[k])


@pytest.mark.parametrize("flux", ("roe", "hllc", "ausmup"))
@pytest.mark.parametrize("name,key", [("synth_p3", "U"), ("quad_q2_p2", "U0"), ("cfg1", "Ur"),
                                      ("tri_p2_q2_backflow", "U")])
def test_vjp_tile_shapes_agree(name, key, flux):
    """The two residual-VJP tile shapes (one-wave tiles of 16 elements, four-wave
    tiles) give the same VJP to round-off, and the autotuner picks one.  The
    one-wave kernel pulls the interface fluxes back with hand-written reverse
    passes; the four-wave kernel differentiates HLLC and AUSM+-up with dual
    numbers, so for them this checks one against the other (the backflow case
    takes HLLC's low-Mach blend)."""
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    bk = backend(ops, flow.replace(flux=flux))
    U = g[key]
    lam = np.random.default_rng(5).standard_normal(U.shape)
    outs = []
    for choice in (0, 1):
        bk.P.vjp_kernel = choice
        outs.append(np.asarray(bk.residual_vjp(U, lam)))
    err = float(np.abs(outs[0] - outs[1]).max() / np.abs(outs[1]).max())
    print(name, "tile shapes differ by", err)
    assert err <= 1e-12
    assert bk.tune_vjp(U, lam) in (0, 1)


@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("name,key", [("cfg1", "Ur"), ("quad_q2_p2", "U0"),
                                      ("tri_p2_q2_backflow", "U")])
def test_fixed_boundary_jacobian_matches_the_pre_pass(name, key, flux):
    """At a fixed U the one-wave VJP applies U's boundary-flux Jacobians
    (``nz_boundary_jacobian``, ``nz_problem.bjac``) to the cotangent's face
    traces instead of running the boundary pre-pass at every call: the same
    VJP to round-off (its traces sum over the face's nodes, the pre-pass's
    over every node), on inflow, outflow, backflow and wall faces."""
    g = golden(name)
    ops, flow = build_case(name, dense=False)
    bk = backend(ops, flow.replace(flux=flux))
    U = g[key]
    lam = np.random.default_rng(9).standard_normal(U.shape)
    bk.P.vjp_kernel = 0
    v0 = np.asarray(bk.residual_vjp(U, lam))
    u, _ = bk._in(U, "state", "U")
    bk._set_boundary_jacobian(bk._p(u))
    try:
        v1 = np.asarray(bk.residual_vjp(U, lam))
    finally:
        bk.P.bjac = None
    err = float(np.abs(v1 - v0).max() / np.abs(v0).max())
    print(name, flux, "fixed Jacobian against the pre-pass", err)
    assert err <= 1e-13
    # the four-wave kernel and the launch with geometry cotangents keep their
    # boundary pre-pass: the Jacobians, when set, leave them unchanged bit
    # for bit
    bk.P.vjp_kernel = 1
    try:
        w0 = np.asarray(bk.residual_vjp(U, lam))
        bk._set_boundary_jacobian(bk._p(u))
        w1 = np.asarray(bk.residual_vjp(U, lam))
    finally:
        bk.P.bjac = None
        bk.P.vjp_kernel = 0
    assert np.array_equal(w0, w1)
    g0 = bk.residual_vjp(U, lam, geometry=True)
    bk._set_boundary_jacobian(bk._p(u))
    try:
        g1 = bk.residual_vjp(U, lam, geometry=True)
    finally:
        bk.P.bjac = None
    for a, b in zip(g0, g1):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("flux", FLUXES)
def test_adjoint_march_graph_replay_matches_eager(flux):
    """The adjoint march's 100-step chunks replayed from a captured HIP graph
    (first chunk eager, second captured, then four replays, each followed by
    the eager defect launch and a host read): the same adjoint state and
    residual bit for bit as the march launched eagerly throughout."""
    g = golden("cfg1")
    ops, flow = build_case("cfg1", dense=False)
    U = g["Ur"]
    seed = np.random.default_rng(17).standard_normal(U.shape) * 1e-3
    out = []
    for graphs in (True, False):
        bk = backend(ops, flow.replace(flux=flux))
        bk.use_graphs = graphs
        lam, rel, its, _ = bk.adjoint_solve(U, seed, tolerance=1e-30, max_iterations=600,
                                            check_interval=100)
        out.append((lam, rel, its))
    (l0, r0, i0), (l1, r1, i1) = out
    print(flux, r0, r1, i0, i1)
    assert i0 == i1 == 600
    assert np.array_equal(l0, l1) and r0 == r1
