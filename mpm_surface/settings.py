"""Surfacing settings per emitter, and ``<cache>/surface_settings.json``, the file the add-on writes.

The file format, defaults, validation messages and the settings fingerprint are the reference's
(src/mpm_surface/settings.py:1-64)::

    {"version": 1, "emitters": {"<emitter name>": {"enabled", "detail", "smoothing", "iso"}}}

An emitter the file does not name is surfaced with the defaults.
"""

from __future__ import annotations

import hashlib
import json
from dataclasses import asdict, dataclass
from pathlib import Path

from mpm_core.cache import write_json

FORMAT_VERSION = 1
FILE_NAME = "surface_settings.json"

# field -> (test a valid value passes, the accepted range as the error message words it)
_RANGES = {
    "detail": (lambda v: v > 0, "> 0"),
    "smoothing": (lambda v: v > 0, "> 0"),
    "iso": (lambda v: 0 < v < 1, "in (0, 1)"),
}


@dataclass(frozen=True)
class SurfaceSettings:
    """``detail``: field cells per particle spacing. ``smoothing``: the kernel radius in units of two particle
    spacings (below about 1 the surface turns lumpy and thin sheets break up). ``iso``: the surface level as a
    fraction of the rest density (0.5 hides sparse material at smoothing 2)."""

    enabled: bool = True
    detail: float = 1.0
    smoothing: float = 2.0
    iso: float = 0.35

    def __post_init__(self):
        for name, (ok, words) in _RANGES.items():
            if not ok(getattr(self, name)):
                raise ValueError(f"{name} must be {words}")

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, d: dict) -> SurfaceSettings:
        return cls(**d)

    def hash(self) -> str:
        """Fingerprint of the settings: SHA-1 of their JSON with sorted keys. It is part of every mesh's key,
        so a mesh made under other settings is made again."""
        text = json.dumps(self.to_dict(), sort_keys=True)
        return hashlib.sha1(text.encode()).hexdigest()


def read_settings(cache_dir: Path) -> dict[str, SurfaceSettings]:
    """The settings the file names, by emitter; {} when the cache has no settings file."""
    path = Path(cache_dir, FILE_NAME)
    if not path.is_file():
        return {}
    doc = json.loads(path.read_text(encoding="utf-8"))
    if doc.get("version") != FORMAT_VERSION:
        raise ValueError(f"unsupported surface settings version {doc.get('version')!r}")
    return {name: SurfaceSettings.from_dict(entry) for name, entry in doc["emitters"].items()}


def write_settings(cache_dir: Path, settings: dict[str, SurfaceSettings]) -> None:
    emitters = {name: s.to_dict() for name, s in settings.items()}
    write_json(Path(cache_dir, FILE_NAME), {"version": FORMAT_VERSION, "emitters": emitters})
