"""Per-emitter surfacing settings and ``<cache>/surface_settings.json``, which the add-on writes
(reference src/mpm_surface/settings.py:1-64; same file format)::

    {"version": 1, "emitters": {"<emitter name>": {"enabled", "detail", "smoothing", "iso"}}}

An emitter absent from the file is surfaced with the defaults.
"""

from __future__ import annotations

import hashlib
import json
from dataclasses import asdict, dataclass
from pathlib import Path

from mpm_core.cache import write_json

FORMAT_VERSION = 1
FILE_NAME = "surface_settings.json"


@dataclass(frozen=True)
class SurfaceSettings:
    """settings.py:21-49. detail: field cells per particle spacing; smoothing: kernel radius in units of two
    particle spacings; iso: the surface level as a fraction of the rest density."""

    enabled: bool = True
    detail: float = 1.0
    smoothing: float = 2.0
    iso: float = 0.35

    def __post_init__(self):
        checks = (("detail", self.detail > 0, "> 0"), ("smoothing", self.smoothing > 0, "> 0"),
                  ("iso", 0 < self.iso < 1, "in (0, 1)"))
        for name, ok, want in checks:
            if not ok:
                raise ValueError(f"{name} must be {want}")

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, d: dict) -> SurfaceSettings:
        return cls(**d)

    def hash(self) -> str:
        """A fingerprint of the settings (part of each mesh's key, so changed settings re-mesh)."""
        return hashlib.sha1(json.dumps(self.to_dict(), sort_keys=True).encode()).hexdigest()


def read_settings(cache_dir: Path) -> dict[str, SurfaceSettings]:
    path = Path(cache_dir) / FILE_NAME
    if not path.exists():
        return {}
    doc = json.loads(path.read_text(encoding="utf-8"))
    if doc.get("version") != FORMAT_VERSION:
        raise ValueError(f"unsupported surface settings version {doc.get('version')!r}")
    return {name: SurfaceSettings.from_dict(entry) for name, entry in doc["emitters"].items()}


def write_settings(cache_dir: Path, settings: dict[str, SurfaceSettings]) -> None:
    write_json(Path(cache_dir) / FILE_NAME,
               {"version": FORMAT_VERSION, "emitters": {name: s.to_dict() for name, s in settings.items()}})
