"""Pseudo-time march to a steady state, and the steady solve around it.

The march has two parts that this module keeps apart:

* :class:`StopRules` decides when a march ends.  It is a small state machine
  fed one residual per convergence check.  Its rules are the reference's
  (``src/timestepping/march.py:125-230``): the first residual must be finite
  and may already meet the tolerance; afterwards, in this order, a
  non-finite residual, the tolerance, growth past ``divergence_factor`` times
  the first residual, and a stall (the last ``stall_window`` checks within
  ``stall_ratio`` of each other) end the march; the iteration budget ends it
  otherwise.  The same object serves the single-design loop below and each
  design of the batched driver (``batch.py``), so both stop on the same
  iteration with the same words.
* :func:`march` only moves the state: one step alone for the initial
  residual, then ``check_interval`` steps per ``backend.run`` call.  With the
  ``hip`` backend each chunk is one replay of a captured HIP graph and the
  state never leaves the device between chunks; the only host work per chunk
  is reading one residual.
"""

from __future__ import annotations

import math
import time
from collections import deque
from collections.abc import Callable
from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np

from .config import SolverOptions, residual_scale


# ---------------------------------------------------------------- results
@dataclass
class SolveHistory:
    """Residual after each convergence check (``march.py:60-70`` field names)."""

    iterations: list[int] = field(default_factory=list)
    residual: list[float] = field(default_factory=list)
    wall_time: float = 0.0

    def as_arrays(self):
        return np.asarray(self.iterations), np.asarray(self.residual)


@dataclass
class SolveResult:
    """Outcome of a steady solve; the field names are the reference's
    (``src/timestepping/march.py:71-122``), so its callers read it unchanged."""

    U: np.ndarray
    operators: object
    flow: object
    converged: bool
    iterations: int
    residual: float
    residual_initial: float
    history: SolveHistory
    backend: str
    residual_scale: float = 1.0
    message: str = ""
    min_density: float = float("nan")
    min_pressure: float = float("nan")
    limiter_activity: float = float("nan")
    mean_repairs: int = 0
    geometry: object = None

    @property
    def residual_scaled(self) -> float:
        """What the tolerance is tested against: residual / (rho_t a_t / L)."""
        return self.residual / self.residual_scale

    @property
    def residual_relative(self) -> float:
        return 0.0 if self.residual_initial == 0.0 else self.residual / self.residual_initial

    def summary(self) -> str:
        head = "converged" if self.converged else "NOT CONVERGED"
        text = (f"{head} after {self.iterations} iterations in {self.history.wall_time:.2f} s "
                f"({self.backend}); scaled residual {self.residual_scaled:.3e}, "
                f"{self.residual_relative:.2e} of the first; "
                f"min rho {self.min_density:.4e}, min p {self.min_pressure:.4e}")
        if self.limiter_activity == self.limiter_activity and self.limiter_activity > 0:
            text += f"; limiter on ~{self.limiter_activity:.2f} elements per call"
        if self.mean_repairs:
            text += f"; {self.mean_repairs} cell means repaired (cfl too large)"
        return text


# ---------------------------------------------------------------- stop rules
class Stop(NamedTuple):
    converged: bool
    reason: str      # "converged", "nonfinite", "diverged", "stalled", "budget"
    message: str     # empty when converged


class StopRules:
    """When a pseudo-time march ends (the reference's rules and order,
    ``src/timestepping/march.py:155-222``).

    ``begin(res0)`` takes the residual of the first step and ``check(done,
    res)`` the residual after each chunk; each returns a :class:`Stop` or
    ``None`` to keep going.  ``activity`` (optional) gives the limiter's mean
    activity when a stall is detected, which picks the advice in the message.
    """

    def __init__(self, opts: SolverOptions, scale: float,
                 activity: Callable[[], float] | None = None):
        self.opts = opts
        self.scale = float(scale)
        self.target = opts.tolerance * self.scale
        self.chunk = min(max(1, opts.check_interval), opts.max_iterations)
        self.activity = activity
        self.res0 = float("nan")
        self.last = float("nan")
        self._recent: deque[float] = deque(maxlen=opts.stall_window)

    # -- the first residual
    def begin(self, res0: float) -> Stop | None:
        self.res0 = self.last = float(res0)
        if not math.isfinite(self.res0):
            return Stop(False, "nonfinite", "initial residual is not finite")
        if self.res0 <= self.target:
            return Stop(True, "converged", "")
        return None

    # -- after each chunk
    def check(self, done: int, res: float) -> Stop | None:
        res = self.last = float(res)
        o = self.opts
        if not math.isfinite(res):
            return Stop(False, "nonfinite",
                        f"residual became non-finite after {done} iterations; reduce cfl "
                        "(try 0.5) or enable limiter='superbee'")
        if res <= self.target:
            return Stop(True, "converged", "")
        if res > o.divergence_factor * self.res0:
            return Stop(False, "diverged",
                        f"residual grew by {res / self.res0:.1e}x after {done} iterations, past "
                        f"divergence_factor={o.divergence_factor:g}; reduce cfl or start from "
                        "initial_condition='quasi1d'")
        self._recent.append(res)
        if len(self._recent) == o.stall_window:
            lo, hi = min(self._recent), max(self._recent)
            if hi > 0.0 and lo / hi > o.stall_ratio:
                return Stop(False, "stalled", self._stall_message(res))
        return None

    def budget(self, done: int) -> Stop:
        """The stop when the iteration budget runs out first."""
        return Stop(False, "budget",
                    f"hit max_iterations={self.opts.max_iterations} at scaled residual "
                    f"{self.last / self.scale:.2e} (target {self.opts.tolerance:g}); raise "
                    "max_iterations or relax tolerance")

    def _stall_message(self, res: float) -> str:
        o = self.opts
        span = o.stall_window * self.chunk
        text = (f"the residual has stalled at {res / self.scale:.3e} (scaled; target "
                f"{o.tolerance:g}) -- it changed by less than {100 * (1 - o.stall_ratio):.0f}% "
                f"over the last {span} iterations")
        act = self.activity() if self.activity is not None else float("nan")
        if math.isfinite(act) and act > 0.0:
            return text + (f", with the limiter active on ~{act:.1f} element(s) per call. That "
                           "is a limit cycle, not slow convergence: the mesh cannot resolve a "
                           "shock or a strong expansion. Refine (refine=+1), reduce the "
                           "polynomial order, or move the operating point")
        return text + ". Try a smaller cfl, scheme='ssprk3', or a finer mesh"


# ---------------------------------------------------------------- the march
def march(U, backend, opts: SolverOptions, scale: float, *, label: str = "",
          progress: Callable[[int, float], None] | None = None):
    """March ``U`` to a steady state with ``backend``.

    Returns ``(U, history, converged, iterations, residual, residual_initial,
    message)``, the reference's tuple (``src/timestepping/march.py:125-230``).
    ``backend`` needs ``run(U, n, scheme) -> (U, residual)``, ``scheme`` and
    ``limiter_activity()``.  ``progress(iterations, residual)`` replaces the
    printed progress lines when given.
    """
    rules = StopRules(opts, scale, activity=backend.limiter_activity)
    hist = SolveHistory()
    scheme = backend.scheme

    U, res0 = backend.run(U, 1, scheme)
    done = 1
    hist.iterations.append(done)
    hist.residual.append(float(res0))
    t_start = time.perf_counter()
    stop = rules.begin(res0)
    printed_at = 0
    while stop is None:
        if done >= opts.max_iterations:
            stop = rules.budget(done)
            break
        n = min(rules.chunk, opts.max_iterations - done)
        U, res = backend.run(U, n, scheme)
        done += n
        hist.iterations.append(done)
        hist.residual.append(float(res))
        if progress is not None:
            progress(done, float(res))
        elif opts.print_interval > 0 and done - printed_at >= opts.print_interval:
            printed_at = done
            print(f"  {label}{done:>7d}   {res / scale:.6e}   ({res / rules.res0:.3e} of initial)")
        stop = rules.check(done, res)
    hist.wall_time = time.perf_counter() - t_start
    return (U, hist, stop.converged, done, float(rules.last), float(rules.res0), stop.message)


def probe_minima(U: np.ndarray, ops, gamma: float) -> tuple[float, float]:
    """Minimum density and pressure over the limiter's probe points: the
    volume points and each element's own face points (``limiter.py:432-445``)."""
    rd = ops.ref
    ed = ops.topology.edges
    vol = np.einsum("iq,eis->eqs", rd.phi_vol, U)
    nface = ops.topology.n_faces
    fb = rd.phi_face[ed.face_side, np.arange(nface)[None, :]]  # (E, nface, N, nqf)
    face = np.einsum("efiq,eis->efqs", fb, U).reshape(U.shape[0], -1, 4)
    pr = np.concatenate([vol, face], axis=1)
    rho = pr[..., 0]
    p = (gamma - 1.0) * (pr[..., 3] - 0.5 * (pr[..., 1] ** 2 + pr[..., 2] ** 2)
                         / np.where(rho != 0.0, rho, 1.0))
    return float(rho.min()), float(p.min())


def solve_steady(ops, flow, opts: SolverOptions | None = None, *, U0, backend: str = "hip",
                 length: float = 1.0, progress=None, geometry=None) -> SolveResult:
    """Steady solve on a prepared operator set from the initial state ``U0``
    (``src/timestepping/march.py:233-299``): march, then refuse a converged
    state that is not physical (the reference's ``not physical`` message)."""
    opts = opts or SolverOptions()
    from . import get_backend

    bk = get_backend(backend, ops, flow, opts)
    U = np.asarray(U0)
    want = (ops.n_elem, ops.ref.n_basis, 4)
    if U.shape != want:
        raise ValueError(f"U0 has shape {U.shape}, expected {want}; change the order of the "
                         "state (initial.change_order) or rebuild the mesh to match")
    scale = residual_scale(flow, length)
    if hasattr(bk, "tune_element_kernel"):
        bk.tune_element_kernel(U)  # large affine meshes: tiles or patch tiles
    U, hist, conv, iters, res, res0, msg = march(bk.asarray(U), bk, opts, scale,
                                                 progress=progress)
    U_np = np.asarray(bk.to_numpy(U))
    rmin, pmin = probe_minima(U_np, ops, flow.gamma)
    if conv and (rmin <= 0.0 or pmin <= 0.0):
        conv = False
        msg = (f"residual converged but the solution is not physical (min rho {rmin:.3e}, "
               f"min p {pmin:.3e}); enable a limiter or refine the mesh")
    return SolveResult(
        U=U_np, operators=ops, flow=flow, converged=conv, iterations=iters, residual=res,
        residual_initial=res0, history=hist, backend=backend, residual_scale=scale, message=msg,
        min_density=rmin, min_pressure=pmin, limiter_activity=bk.limiter_activity(),
        mean_repairs=int(bk.n_mean_repaired), geometry=geometry,
    )
