"""Output formats of solves and sweeps.

* :func:`save_solution` writes a solve the way the reference's ``solve
  --save`` does (``src/cli.py:166-174``): ``U``, ``converged``, ``residual``
  and one ``geom_<field>`` entry per geometry field that is set, in a
  compressed ``.npz``.
* :class:`SweepResult` is the reference's sweep table (``src/sweep.py:68-145``):
  the swept parameters of every point, the metrics ``METRICS``
  (``src/sweep.py:45-65``), the convergence flags, the operating regimes and
  the solver messages, with ``table()``, ``to_csv()`` and ``to_dict()``.
  :func:`sweep_metrics` fills one row from a design's state, with the
  performance integrals computed on the device.
"""

from __future__ import annotations

import dataclasses
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

# the reference's sweep metrics, in its column order (src/sweep.py:45-65)
METRICS = (
    "thrust", "thrust_wall", "thrust_coefficient", "thrust_imbalance", "thrust_efficiency",
    "specific_thrust", "mass_flow_in", "mass_imbalance", "entropy_error", "exit_mach",
    "exit_pressure_ratio", "ideal_thrust", "quasi1d_exit_mach", "quasi1d_shock_x", "iterations",
    "wall_time", "residual", "min_density", "min_pressure",
)


def save_solution(path: str, result, geometry=None) -> None:
    """``result.U``, ``converged`` and ``residual`` plus ``geom_<field>`` for
    every geometry field that is not None (``src/cli.py:166-174``)."""
    geom = geometry if geometry is not None else getattr(result, "geometry", None)
    fields = {} if geom is None else {f"geom_{k}": v for k, v in dataclasses.asdict(geom).items()
                                      if v is not None}
    np.savez_compressed(path, U=np.asarray(result.U), converged=bool(result.converged),
                        residual=float(result.residual), **fields)


@dataclass
class SweepResult:
    """Tabulated results of a parameter sweep (``src/sweep.py:68-145``):
    ``result['thrust']`` or ``result['area_ratio']`` give arrays aligned with
    ``points``."""

    parameters: tuple[str, ...]
    points: list[dict[str, float]]
    values: dict[str, np.ndarray]
    converged: np.ndarray
    regimes: list[str]
    messages: list[str]
    shape: tuple[int, ...] = ()

    def __getitem__(self, key: str) -> np.ndarray:
        if key in self.values:
            return self.values[key]
        if key in self.parameters:
            return np.asarray([p[key] for p in self.points])
        raise KeyError(f"{key!r} is neither a swept parameter {self.parameters} nor a metric; "
                       f"metrics are {tuple(self.values)}")

    @property
    def n_points(self) -> int:
        return len(self.points)

    @property
    def all_converged(self) -> bool:
        return bool(np.all(self.converged))

    def reshape(self, key: str) -> np.ndarray:
        return self[key].reshape(self.shape)

    def failures(self) -> list[tuple[dict[str, float], str]]:
        return [(p, m) for p, m, ok in zip(self.points, self.messages, self.converged) if not ok]

    def table(self, metrics: Sequence[str] = ("thrust_coefficient", "exit_mach",
                                              "entropy_error")) -> str:
        """Fixed-width text: the parameters, the chosen metrics, converged."""
        cols = list(self.parameters) + list(metrics) + ["ok"]
        head = "  ".join(f"{c:>18s}" for c in cols)
        lines = [head, "-" * len(head)]
        for i, pt in enumerate(self.points):
            cells = [f"{pt[p]:18.6g}" for p in self.parameters]
            cells += [f"{self.values[m][i]:18.6g}" for m in metrics]
            cells.append(f"{'yes' if self.converged[i] else 'NO':>18s}")
            lines.append("  ".join(cells))
        return "\n".join(lines)

    def to_csv(self, path: str) -> None:
        """Every parameter and metric, then converged and regime, one row per
        point; floats written with ``repr`` so they read back exactly."""
        cols = list(self.parameters) + list(self.values) + ["converged", "regime"]
        with open(path, "w", encoding="utf-8") as fh:
            fh.write(",".join(cols) + "\n")
            for i, pt in enumerate(self.points):
                cells = [repr(pt[p]) for p in self.parameters]
                cells += [repr(float(self.values[m][i])) for m in self.values]
                cells += [str(bool(self.converged[i])), self.regimes[i]]
                fh.write(",".join(cells) + "\n")

    def to_dict(self) -> dict[str, np.ndarray]:
        out = {p: self[p] for p in self.parameters}
        out.update(self.values)
        out["converged"] = self.converged
        return out


def read_csv(path: str) -> dict[str, list[str]]:
    """Columns of a sweep CSV as strings (for round trips and tools)."""
    with open(path, encoding="utf-8") as fh:
        rows = [ln.rstrip("\n").split(",") for ln in fh]
    head = rows[0]
    return {c: [r[k] for r in rows[1:]] for k, c in enumerate(head)}


def sweep_metrics(perf, result, geom, flow, quasi1d) -> dict[str, float]:
    """One sweep row (``_extract``, ``src/sweep.py:192-215``) from a design's
    ``Performance``, its solve record and its quasi-1D solution."""
    return {
        "thrust": perf.thrust, "thrust_wall": perf.thrust_wall,
        "thrust_coefficient": perf.thrust_coefficient, "thrust_imbalance": perf.thrust_imbalance,
        "thrust_efficiency": perf.thrust_efficiency, "specific_thrust": perf.specific_thrust,
        "mass_flow_in": perf.mass_flow_in, "mass_imbalance": perf.mass_imbalance,
        "entropy_error": perf.entropy_error, "exit_mach": perf.exit_mach_area_averaged,
        "exit_pressure_ratio": perf.exit_pressure_ratio, "ideal_thrust": perf.ideal_thrust,
        "quasi1d_exit_mach": quasi1d.exit_mach,
        "quasi1d_shock_x": float("nan") if quasi1d.shock_x is None else quasi1d.shock_x,
        "iterations": float(result.iterations),
        "wall_time": float(getattr(result, "wall_time", float("nan"))),
        "residual": float(result.residual),
        "min_density": float(getattr(result, "min_density", float("nan"))),
        "min_pressure": float(getattr(result, "min_pressure", float("nan"))),
    }
