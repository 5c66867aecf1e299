"""Per-step result record (upstream steps/_runner.py:26-47)."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List


@dataclass
class StepResult:
    name: str
    type: str
    cycle: int
    t_start_cycle: float
    duration_planned: float
    duration: float
    stopped_early: bool
    acc: Dict[str, float]
    inv_start: Dict[str, float]
    inv_end: Dict[str, float]
    outlet: List[Dict] = field(default_factory=list)
    snapshots: List[Dict] = field(default_factory=list)
    solver: Dict = field(default_factory=dict)
    wall_time: float = 0.0
    T_jacket: float = 298.15
    jacket_on: bool = True
    product_port: str = "V3"
    valves: Dict = field(default_factory=dict)
    ads_start: Dict[str, float] = field(default_factory=dict)
    ads_end: Dict[str, float] = field(default_factory=dict)
