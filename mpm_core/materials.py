"""Material parameters and presets.

Interface follows the reference's ``mpm_core.materials`` (src/mpm_core/materials.py:1-158):
``MaterialParams`` with the same fields, validation and ``lame()``; presets by id with
``list_presets`` / ``load_preset`` / ``save_user_preset`` / ``delete_user_preset``.

Built-in presets live in one table, ``presets/builtin.toml``. User presets are single TOML
files (``label``, optional ``description``, ``[material]``, ``[look]``) in a folder the caller
names, with ids ``user:<file stem>``; that is the file format the Blender add-on writes, so it is
kept as the reference defines it.
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


@dataclass(frozen=True)
class MaterialParams:
    """One material (src/mpm_core/materials.py:24-59)."""

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
        if not self.youngs_modulus > 0:
            raise ValueError("youngs_modulus must be > 0")
        if not 0 <= self.poisson_ratio < 0.5:
            raise ValueError("poisson_ratio must be in [0, 0.5)")
        if not self.density > 0:
            raise ValueError("density must be > 0")

    def lame(self) -> tuple[float, float]:
        """Lamé parameters (mu, lam) in double precision (materials.py:48-51)."""
        e, nu = self.youngs_modulus, self.poisson_ratio
        mu = e / (2.0 * (1.0 + nu))
        lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        return mu, lam

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, d: dict) -> MaterialParams:
        names = {f.name for f in fields(cls)}
        return cls(**{k: v for k, v in d.items() if k in names})


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


def _builtin_table() -> dict[str, dict]:
    with open(_BUILTIN_FILE, "rb") as f:
        return _toml.load(f)


def _user_files(user_dir: Path | None) -> dict[str, Path]:
    if user_dir is None or not Path(user_dir).is_dir():
        return {}
    return {USER_PREFIX + p.stem: p for p in Path(user_dir).glob("*.toml")}


def _preset_from_doc(preset_id: str, doc: dict) -> Preset:
    look = doc["look"]
    return Preset(
        id=preset_id,
        label=doc["label"],
        material=MaterialParams(name=preset_id, **doc["material"]),
        look=Look(look["display"], tuple(look["base_color"]), look["roughness"], look["subsurface"]),
        description=doc.get("description", ""),
    )


def list_presets(user_dir: Path | None = None) -> list[str]:
    """Built-in ids sorted, then user ids sorted by label (materials.py:98-101)."""
    users = list(_user_files(user_dir))
    users.sort(key=lambda pid: load_preset(pid, user_dir).label.lower())
    return sorted(_builtin_table()) + users


def load_preset(preset_id: str, user_dir: Path | None = None) -> Preset:
    builtin = _builtin_table()
    if preset_id in builtin:
        return _preset_from_doc(preset_id, builtin[preset_id])
    user = _user_files(user_dir)
    if preset_id in user:
        return _preset_from_doc(preset_id, _toml.loads(user[preset_id].read_text(encoding="utf-8")))
    raise KeyError(f"unknown preset {preset_id!r}, available: {sorted([*builtin, *user])}")


def user_preset_id(label: str) -> str:
    """Id for a user preset label: lower case, runs of other characters folded to ``_``."""
    stem = re.sub(r"[^a-z0-9]+", "_", label.strip().lower()).strip("_")
    if not stem:
        raise ValueError("a preset name needs at least one letter or digit")
    return USER_PREFIX + stem


def save_user_preset(user_dir: Path, label: str, material: MaterialParams, look: Look,
                     description: str = "") -> Preset:
    """Write (or replace) a user preset; the label of a built-in preset is refused."""
    label = label.strip()
    builtin_labels = {doc["label"].lower() for doc in _builtin_table().values()}
    if label.lower() in builtin_labels:
        raise ValueError(f"{label!r} is the name of a built-in preset")
    preset_id = user_preset_id(label)
    out = [f"label = {json.dumps(label)}"]
    if description:
        out.append(f"description = {json.dumps(description)}")
    out += ["", "[material]"]
    out += [f"{k} = {json.dumps(v)}" for k, v in material.to_dict().items() if k != "name"]
    out += ["", "[look]", f"display = {json.dumps(look.display)}",
            f"base_color = {json.dumps([float(c) for c in look.base_color])}",
            f"roughness = {json.dumps(float(look.roughness))}",
            f"subsurface = {json.dumps(float(look.subsurface))}", ""]
    folder = Path(user_dir)
    folder.mkdir(parents=True, exist_ok=True)
    path = folder / (preset_id[len(USER_PREFIX):] + ".toml")
    tmp = path.with_name(path.name + ".tmp")
    tmp.write_text("\n".join(out), encoding="utf-8")
    os.replace(tmp, path)
    return load_preset(preset_id, folder)


def delete_user_preset(user_dir: Path, preset_id: str) -> None:
    if not preset_id.startswith(USER_PREFIX):
        raise ValueError(f"{preset_id!r} is a built-in preset")
    path = _user_files(user_dir).get(preset_id)
    if path is None:
        raise KeyError(f"unknown user preset {preset_id!r}")
    path.unlink()
