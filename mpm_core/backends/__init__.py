"""Solver backend registry (src/mpm_core/backends/__init__.py:9-28).

Backends are imported on first use, so importing this module stays cheap. ``"hip"`` is the
MI355X backend of this package; the reference's ``"warp"`` backend is not shipped here.
"""

from __future__ import annotations

from collections.abc import Callable

from mpm_core.solver import Solver

_factories: dict[str, Callable[..., Solver]] = {}


def _hip_factory(**kwargs) -> Solver:
    from mpm_core.backends.hip.solver import HipSolver

    return HipSolver(**kwargs)


_factories["hip"] = _hip_factory


def register_backend(name: str, factory: Callable[..., Solver]) -> None:
    _factories[name] = factory


def create_solver(backend: str, **kwargs) -> Solver:
    try:
        factory = _factories[backend]
    except KeyError:
        raise KeyError(f"unknown solver backend {backend!r}, known: {sorted(_factories)}") from None
    return factory(**kwargs)
