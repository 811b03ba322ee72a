"""Phase timing of an alignment, the schema of cugal/profile.py:14-97
(Phase, SinkhornProfile, TimeStamp, Profile with to_dict/from_dict/average).
The reference's matplotlib/CSV helpers (cugal/profile.py:100-172) are not on
the alignment path and are not provided.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from enum import Enum
from time import time

import numpy as np
import torch


class Phase(Enum):
    FEATURE_EXTRACTION = "FEATURE_EXTRACTION"
    SINKHORN = "SINKHORN"
    HUNGARIAN = "HUNGARIAN"
    GRADIENT = "GRADIENT"


@dataclass
class SinkhornProfile:
    errors: list[float] = field(default_factory=list)
    iteration_count: int = 0
    time: float = 0.0

    @staticmethod
    def average(profiles: list["SinkhornProfile"]) -> "SinkhornProfile":
        time_ = sum(p.time for p in profiles) / len(profiles)
        iteration_count = int(sum(p.iteration_count for p in profiles) / len(profiles))
        return SinkhornProfile([], iteration_count, time_)


class TimeStamp:
    """A point in time: a device event on a ROCm device (recorded on the
    current stream), else the host clock (cugal/profile.py:33-50)."""

    def __init__(self, device: str):
        if "cuda" in device:
            self.time = torch.cuda.Event(enable_timing=True)
            self.time.record()
        else:
            self.time = time()
        self.device = device

    def is_cuda_event(self) -> bool:
        return isinstance(self.time, torch.cuda.Event)

    def elapsed_seconds(self, earlier: "TimeStamp") -> float:
        if self.is_cuda_event():
            torch.cuda.synchronize()
            return earlier.time.elapsed_time(self.time) / 1000
        return self.time - earlier.time


@dataclass
class Profile:
    sinkhorn_profiles: list[SinkhornProfile] = field(default_factory=list)
    phase_times: dict[Phase, float] = field(default_factory=dict)
    time: float = 0.0
    max_memory: int | None = None
    frank_wolfe_iterations: int = 0
    sparsity: float = 0.0

    def log_time(self, start_time: TimeStamp, phase: Phase):
        now = TimeStamp(start_time.device)
        prev_time = self.phase_times.get(phase, 0)
        self.phase_times[phase] = prev_time + now.elapsed_seconds(start_time)

    def to_dict(self) -> dict:
        d = dataclasses.asdict(self)
        d["sinkhorn_profiles"] = [dataclasses.asdict(p) for p in self.sinkhorn_profiles]
        d["phase_times"] = {phase.value: t for phase, t in self.phase_times.items()}
        return d

    @classmethod
    def from_dict(cls, d: dict):
        d = dict(d)
        d["sinkhorn_profiles"] = [SinkhornProfile(**p) for p in d["sinkhorn_profiles"]]
        d["phase_times"] = {Phase[phase]: t for phase, t in d["phase_times"].items()}
        return cls(**d)

    @staticmethod
    def average(profiles: list["Profile"]) -> "Profile":
        phase_times = {
            phase: sum(p.phase_times[phase] for p in profiles) / len(profiles)
            for phase in profiles[0].phase_times.keys()
        }
        time_ = sum(p.time for p in profiles) / len(profiles)
        max_memory = max([p.max_memory for p in profiles if p.max_memory is not None], default=0)
        sinkhorn_count = len(profiles[0].sinkhorn_profiles)
        frank_wolfe_iterations = np.mean([p.frank_wolfe_iterations for p in profiles])
        sparsity = np.mean([p.sparsity for p in profiles])
        if all(len(p.sinkhorn_profiles) == sinkhorn_count for p in profiles):
            sinkhorn_profiles = [
                SinkhornProfile.average([p.sinkhorn_profiles[i] for p in profiles]) for i in range(sinkhorn_count)
            ]
        else:
            sinkhorn_profiles = []
        return Profile(sinkhorn_profiles, phase_times, time_, max_memory, frank_wolfe_iterations, sparsity)
