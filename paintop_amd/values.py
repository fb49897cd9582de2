"""Host-side resource values in paintop's layout
(paintop-core/src/executor/value.rs:35-85; descriptors paintop-ir/src/resource.rs).

A value owns a flat row-major, channel-interleaved f32 buffer whose length is
``width * height * channels``, validated on construction like
``ResourceValue::new``. A report value has no raster.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# ChannelLayout -> channel_count (paintop-ir/src/resource.rs:196-216)
CHANNEL_COUNTS = {"Gray": 1, "GrayAlpha": 2, "Rgb": 3, "Rgba": 4}


@dataclass(frozen=True)
class Extent:
    width: int
    height: int


@dataclass(frozen=True)
class ImageDescriptor:
    extent: Extent
    layout: str = "Gray"  # Gray | GrayAlpha | Rgb | Rgba
    color: str = "LinearSrgb"  # ColorEncoding: LinearSrgb | Srgb (resource.rs)

    @property
    def channels(self) -> int:
        return CHANNEL_COUNTS[self.layout]


@dataclass(frozen=True)
class MaskDescriptor:
    extent: Extent


@dataclass
class SolverData:
    """paintop_ir::SolverData (resource.rs:1130-1176), iterative fields filled
    as ``solver_data`` does (poisson_blend.rs:249-266)."""

    kind: str
    steps: int
    stability_number: float
    stability_limit: float
    stable: bool
    residual_history: list
    total_energy: float
    iterations: Optional[int]
    stop_reason: Optional[str]  # "converged" | "max-iterations" | "stalled"
    converged: Optional[bool]
    tolerance: Optional[float]
    final_residual: Optional[float]

    def to_json(self) -> dict:
        return {
            "kind": self.kind,
            "steps": self.steps,
            "stability_number": self.stability_number,
            "stability_limit": self.stability_limit,
            "stable": self.stable,
            "residual_history": list(self.residual_history),
            "total_energy": self.total_energy,
            "iterations": self.iterations,
            "stop_reason": self.stop_reason,
            "converged": self.converged,
            "tolerance": self.tolerance,
            "final_residual": self.final_residual,
        }


@dataclass
class AssertionOutcome:
    """paintop_ir::AssertionOutcome (resource.rs:982-1040), the fields
    assert.no_change_outside_mask@1 fills (assert.rs:694-708)."""

    assertion: str
    passed: bool
    severity: str  # error | warning | metric
    max_abs_delta_outside: Optional[float] = None
    changed_pixels_outside: Optional[int] = None
    worst_pixel: Optional[tuple] = None
    locations: list = field(default_factory=list)

    def fails_run(self) -> bool:
        """AssertionSeverity::fails_run (resource.rs:962-964): only an
        `error` severity fails the run."""
        return (not self.passed) and self.severity == "error"


@dataclass
class Report:
    """The op report (resource.rs:1066-1113) as the Poisson ops fill it
    (poisson_blend.rs:581-593), or an assertion's verdict."""

    extent: Extent
    channels: int
    all_finite: bool
    solver: Optional[SolverData]
    channel_stats: list = field(default_factory=list)
    content_hash: str = ""
    assertion: Optional[AssertionOutcome] = None


class ResourceValue:
    """A descriptor plus its samples (value.rs:35-85)."""

    def __init__(self, descriptor, channels: int, samples=None, report: Optional[Report] = None):
        self.descriptor = descriptor
        self.channels = channels
        self.report = report
        if report is not None:
            self.samples = np.zeros(0, dtype=np.float32)
            return
        arr = np.ascontiguousarray(np.asarray(samples, dtype=np.float32).reshape(-1))
        ext = descriptor.extent
        expected = ext.width * ext.height * channels
        if arr.size != expected:
            from .errors import ErrorClass, PaintopError
            raise PaintopError(ErrorClass.EXECUTION, "E_RESOURCE_BUFFER",
                               f"sample buffer length {arr.size} != {expected}")
        self.samples = arr

    @classmethod
    def report_value(cls, report: Report) -> "ResourceValue":
        return cls(None, report.channels, report=report)

    def extent(self) -> Extent:
        return self.descriptor.extent

    def as_report(self) -> Optional[Report]:
        return self.report

    def array(self) -> np.ndarray:
        """The samples as (height, width, channels)."""
        e = self.descriptor.extent
        return self.samples.reshape(e.height, e.width, self.channels)


def image(width: int, height: int, samples, layout: str = "Gray") -> ResourceValue:
    d = ImageDescriptor(Extent(width, height), layout)
    return ResourceValue(d, d.channels, samples)


def mask(width: int, height: int, samples) -> ResourceValue:
    return ResourceValue(MaskDescriptor(Extent(width, height)), 1, samples)
