"""Materials and their presets.

The interface is the reference's ``mpm_core.materials`` (src/mpm_core/materials.py:24-158): ``MaterialParams``
with the same fields, defaults, validation messages and ``lame()``; a ``Look`` and a ``Preset``; presets by id
through ``list_presets`` / ``load_preset`` / ``user_preset_id`` / ``save_user_preset`` / ``delete_user_preset``.

Built-in presets are one table, ``presets/builtin.toml`` (one section per preset id). A user preset is a TOML
file in a folder the caller names: top-level ``label`` and optional ``description``, then ``[material]`` and
``[look]`` sections; its id is ``user:<file stem>``. The Blender add-on writes and reads that file format, so
it is kept as the reference defines it.
"""

from __future__ import annotations

import json
import os
import re
from dataclasses import asdict, dataclass, fields
from pathlib import Path

try:  # Python >= 3.11
    import tomllib as _toml
except ModuleNotFoundError:  # Python 3.10: the same parser under its package name
    import tomli as _toml

MODELS = ("elastic", "von_mises", "drucker_prager", "snow")
DISPLAYS = ("mesh", "points")
USER_PREFIX = "user:"
_BUILTIN_FILE = Path(__file__).with_name("presets") / "builtin.toml"

# (condition a valid material meets, message when it does not), checked after the model name
_RULES = (
    (lambda m: m.youngs_modulus > 0, "youngs_modulus must be > 0"),
    (lambda m: 0 <= m.poisson_ratio < 0.5, "poisson_ratio must be in [0, 0.5)"),
    (lambda m: m.density > 0, "density must be > 0"),
)


@dataclass(frozen=True)
class MaterialParams:
    """One material: the constitutive model and its parameters (materials.py:24-59). ``yield_stress`` is von
    Mises's, ``friction_angle_deg`` Drucker-Prager's, the last four snow's."""

    name: str
    model: str
    youngs_modulus: float
    poisson_ratio: float
    density: float
    yield_stress: float = 0.0
    friction_angle_deg: float = 0.0
    hardening: float = 0.0
    critical_compression: float = 0.0
    critical_stretch: float = 0.0
    max_hardening: float = 5.0

    def __post_init__(self):
        if self.model not in MODELS:
            raise ValueError(f"unknown material model {self.model!r}, expected one of {MODELS}")
        for holds, message in _RULES:
            if not holds(self):
                raise ValueError(message)

    def lame(self) -> tuple[float, float]:
        """(mu, lambda) from Young's modulus E and Poisson's ratio nu, in double precision:
        mu = E / (2 (1 + nu)), lambda = E nu / ((1 + nu)(1 - 2 nu))."""
        e, nu = float(self.youngs_modulus), float(self.poisson_ratio)
        return e / (2.0 * (1.0 + nu)), e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, d: dict) -> MaterialParams:
        """Build from a dict, ignoring keys that are not fields (documents may carry extra ones)."""
        known = {f.name for f in fields(cls)}
        return cls(**{k: v for k, v in d.items() if k in known})


@dataclass(frozen=True)
class Look:
    display: str
    base_color: tuple[float, float, float]
    roughness: float
    subsurface: float

    def __post_init__(self):
        if self.display not in DISPLAYS:
            raise ValueError(f"unknown display {self.display!r}, expected one of {DISPLAYS}")


@dataclass(frozen=True)
class Preset:
    id: str
    label: str
    material: MaterialParams
    look: Look
    description: str = ""

    @property
    def is_user(self) -> bool:
        return self.id.startswith(USER_PREFIX)


# ---- preset documents ----------------------------------------------------------------------------------

def _preset(preset_id: str, doc: dict) -> Preset:
    look = doc["look"]
    return Preset(id=preset_id, label=doc["label"], material=MaterialParams(name=preset_id, **doc["material"]),
                  look=Look(look["display"], tuple(look["base_color"]), look["roughness"], look["subsurface"]),
                  description=doc.get("description", ""))


def _toml_text(label: str, description: str, material: MaterialParams, look: Look) -> str:
    """A user preset file. Every value is a string, a number, a boolean or a list of numbers, whose JSON
    spelling is also valid TOML."""
    head = {"label": label}
    if description:
        head["description"] = description
    sections = {
        "material": {k: v for k, v in material.to_dict().items() if k != "name"},
        "look": {"display": look.display, "base_color": [float(c) for c in look.base_color],
                 "roughness": float(look.roughness), "subsurface": float(look.subsurface)},
    }
    lines = [f"{k} = {json.dumps(v)}" for k, v in head.items()]
    for name, table in sections.items():
        lines += ["", f"[{name}]"] + [f"{k} = {json.dumps(v)}" for k, v in table.items()]
    return "\n".join(lines + [""])


def _builtin() -> dict[str, dict]:
    with open(_BUILTIN_FILE, "rb") as f:
        return _toml.load(f)


def _user_paths(user_dir: Path | None) -> dict[str, Path]:
    folder = None if user_dir is None else Path(user_dir)
    if folder is None or not folder.is_dir():
        return {}
    return {USER_PREFIX + p.stem: p for p in folder.glob("*.toml")}


def list_presets(user_dir: Path | None = None) -> list[str]:
    """Built-in ids in sorted order, then the user presets of ``user_dir`` ordered by label, case-insensitive
    (materials.py:98-101)."""
    labels = {pid: load_preset(pid, user_dir).label.lower() for pid in _user_paths(user_dir)}
    return sorted(_builtin()) + sorted(labels, key=labels.__getitem__)


def load_preset(preset_id: str, user_dir: Path | None = None) -> Preset:
    builtin = _builtin()
    doc = builtin.get(preset_id)
    if doc is None:
        users = _user_paths(user_dir)
        if preset_id not in users:
            raise KeyError(f"unknown preset {preset_id!r}, available: {sorted(list(builtin) + list(users))}")
        doc = _toml.loads(users[preset_id].read_text(encoding="utf-8"))
    return _preset(preset_id, doc)


def user_preset_id(label: str) -> str:
    """The id of a user preset labelled ``label``: lower case, every run of characters other than a-z and 0-9
    folded to one ``_``, outer ones dropped (labels that differ only in case or punctuation share an id)."""
    stem = "_".join(part for part in re.split(r"[^a-z0-9]+", label.strip().lower()) if part)
    if not stem:
        raise ValueError("a preset name needs at least one letter or digit")
    return USER_PREFIX + stem


def save_user_preset(user_dir: Path, label: str, material: MaterialParams, look: Look,
                     description: str = "") -> Preset:
    """Write a user preset, replacing the one with the same id; a built-in preset's label is refused."""
    label = label.strip()
    if any(doc["label"].lower() == label.lower() for doc in _builtin().values()):
        raise ValueError(f"{label!r} is the name of a built-in preset")
    preset_id = user_preset_id(label)
    folder = Path(user_dir)
    folder.mkdir(parents=True, exist_ok=True)
    path = folder / f"{preset_id[len(USER_PREFIX):]}.toml"
    tmp = folder / f"{path.name}.tmp"
    tmp.write_text(_toml_text(label, description, material, look), encoding="utf-8")
    os.replace(tmp, path)
    return load_preset(preset_id, folder)


def delete_user_preset(user_dir: Path, preset_id: str) -> None:
    if not preset_id.startswith(USER_PREFIX):
        raise ValueError(f"{preset_id!r} is a built-in preset")
    path = _user_paths(user_dir).get(preset_id)
    if path is None:
        raise KeyError(f"unknown user preset {preset_id!r}")
    path.unlink()
