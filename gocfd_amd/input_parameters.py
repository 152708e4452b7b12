"""InputParameters2D: the YAML solver parameters of the reference.

Mirrors InputParameters/InputParameters.go:11-32 and the CLI defaults of
cmd/2D.go:86-88 (Gamma 1.4, Minf 0.1). The reference parses with
ghodss/yaml, which converts YAML to JSON and binds keys to the Go field
names case-insensitively (the struct's yaml tags are not used), so a key
"LocalTimeStepping" sets LocalTimeStepping while "LocalTimeStep" does not.
Unknown keys are ignored, as encoding/json ignores them.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Any, Dict, List

import yaml

# Field names of InputParameters2D in declaration order
_FIELDS = ["Title", "CFL", "FluxType", "InitType", "PolynomialOrder", "FinalTime", "Minf",
           "Gamma", "Alpha", "BCs", "LocalTimeStepping", "MaxIterations", "ImplicitSolver",
           "Limiter", "Kappa", "PlotFields"]


@dataclass
class InputParameters2D:
    Title: str = ""
    CFL: float = 0.0
    FluxType: str = ""
    InitType: str = ""
    PolynomialOrder: int = 0
    FinalTime: float = 0.0
    Minf: float = 0.1
    Gamma: float = 1.4
    Alpha: float = 0.0
    BCs: Dict[str, Any] = field(default_factory=dict)
    LocalTimeStepping: bool = False
    MaxIterations: int = 0
    ImplicitSolver: bool = False
    Limiter: str = ""
    Kappa: float = 0.0
    PlotFields: List[str] = field(default_factory=list)

    def parse(self, data: str | bytes) -> "InputParameters2D":
        """InputParameters2D.Parse (InputParameters.go:30-32)."""
        doc = yaml.safe_load(data) or {}
        if not isinstance(doc, dict):
            raise ValueError("input parameters must be a YAML mapping")
        lower = {f.lower(): f for f in _FIELDS}
        for key, val in doc.items():
            name = lower.get(str(key).lower())
            if name is None:
                continue
            cur = getattr(self, name)
            if isinstance(cur, bool):
                val = bool(val)
            elif isinstance(cur, int):
                val = int(val)
            elif isinstance(cur, float):
                val = float(val)
            elif isinstance(cur, str):
                val = "" if val is None else str(val)
            elif isinstance(cur, list):
                val = [str(v) for v in (val or [])]
            setattr(self, name, val)
        return self

    @classmethod
    def from_file(cls, path: str) -> "InputParameters2D":
        with open(path, "rb") as f:
            return cls().parse(f.read())

    def to_dict(self):
        return dataclasses.asdict(self)
