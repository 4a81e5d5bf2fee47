"""Project files: the case-JSON input boundary.

Accepts the reference's `aerokinetic_project` schema v1
(`src/aerokinetic/project.py:38-92`): embedded polygons, `.dat` files, or a
`naca4` / `naca0012` generator dict. `case_fingerprint` hashes only geometry,
flow and grid exactly as the reference does, so a saved reference project and
this package agree on the fingerprint. The refined schema
(`aerokinetic_refined_project`, `src/aerokinetic/refined_solver.py:102-112`)
is loaded by `load_refined_project` in `refined.py`.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass, field
import hashlib
import json
import math
from pathlib import Path

from ._io import write_json
from .config import Atmosphere, Flow, Grid, Plan
from .convergence import StopCriteria
from .geometry import Contour, Section, naca0012, naca4


@dataclass(frozen=True)
class RunSettings:
    duration_convective: float = 0.2
    sample_every: int = 10
    frame_every: int = 20
    check_every: int = 10
    average_after: float = 0.1
    checkpoint_every: int = 1000
    termination: StopCriteria = field(default_factory=StopCriteria)

    def __post_init__(self):
        if isinstance(self.termination, dict):
            object.__setattr__(self, "termination", StopCriteria(**self.termination))
        if not isinstance(self.termination, StopCriteria):
            raise ValueError("termination must contain stop criteria")
        if not math.isfinite(self.duration_convective) or self.duration_convective <= 0:
            raise ValueError("duration_convective must be finite and positive")
        if not math.isfinite(self.average_after) or self.average_after < 0:
            raise ValueError("average_after must be finite and nonnegative")
        for name in ("sample_every", "frame_every", "check_every", "checkpoint_every"):
            v = getattr(self, name)
            lo = 0 if name == "checkpoint_every" else 1
            if type(v) is not int or v < lo:
                raise ValueError(f"{name} has to be an integer >= {lo}")


def _elements(geometry, base):
    if "elements" in geometry:
        out = []
        for i, item in enumerate(geometry["elements"]):
            if "file" in item:
                out.append(Contour.from_dat(Path(base) / item["file"]))
            else:
                out.append(Contour(item["xy"], item.get("name", f"element_{i}")))
        return tuple(out)
    if "naca4" in geometry:
        return (naca4(**geometry["naca4"]),)
    return (naca0012(**geometry.get("naca0012", {})),)


@dataclass(frozen=True)
class Project:
    section: Section
    flow: Flow = field(default_factory=Flow)
    grid: Grid = field(default_factory=Grid)
    run: RunSettings = field(default_factory=RunSettings)
    title: str = "NACA0012"

    @classmethod
    def from_dict(cls, data, base_directory=Path(".")):
        if data.get("format", "aerokinetic_project") != "aerokinetic_project" or data.get("schema_version", 1) != 1:
            raise ValueError("Unsupported project format or schema version")
        geometry = data.get("geometry", {})
        elements = _elements(geometry, base_directory)
        flow = dict(data.get("flow", {}))
        flow["atmosphere"] = Atmosphere(**flow.get("atmosphere", {}))
        section = Section(elements, geometry.get("chord"), geometry.get("moment_reference"))
        return cls(section, Flow(**flow), Grid(**data.get("grid", {})),
                   RunSettings(**data.get("run", {})), data.get("title", elements[0].name))

    @classmethod
    def load(cls, path):
        path = Path(path)
        return cls.from_dict(json.loads(path.read_text(encoding="utf-8-sig")), path.resolve().parent)

    def as_dict(self):
        grid = asdict(self.grid)
        if grid["farfield"] == "legacy":
            # Older projects have no farfield key; omitting the default keeps
            # their fingerprints unchanged.
            del grid["farfield"]
        geometry = {"elements": [{"name": e.name, "xy": e.xy.tolist()} for e in self.section.elements],
                    "chord": self.section.chord, "moment_reference": self.section.moment_reference}
        return {"format": "aerokinetic_project", "schema_version": 1, "title": self.title,
                "geometry": geometry, "flow": asdict(self.flow), "grid": grid, "run": asdict(self.run)}

    def save(self, path):
        write_json(path, self.as_dict())

    def plan(self):
        return Plan.create(self.section, self.flow, self.grid)

    @property
    def case_fingerprint(self):
        data = self.as_dict()
        case = {k: data[k] for k in ("geometry", "flow", "grid")}
        text = json.dumps(case, sort_keys=True, separators=(",", ":"))
        return hashlib.sha256(text.encode()).hexdigest()
