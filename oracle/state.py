"""Reconstruction state containers (reference phaser/state.py:17-211), numpy only.
Test infrastructure only."""
import copy
from dataclasses import dataclass, field
import typing as t

import numpy


@dataclass
class Patterns:
    patterns: numpy.ndarray       # corner-origin diffraction intensities
    pattern_mask: numpy.ndarray


@dataclass
class IterState:
    engine_num: int = 0
    engine_iter: int = 0
    total_iter: int = 0
    n_engine_iters: t.Optional[int] = None
    n_total_iters: t.Optional[int] = None


@dataclass
class ProbeState:
    sampling: t.Any               # oracle.num.Sampling
    data: numpy.ndarray           # (modes, y, x)


@dataclass
class ObjectState:
    sampling: t.Any               # oracle.object.ObjectSampling
    data: numpy.ndarray           # (z, y, x)
    thicknesses: numpy.ndarray


@dataclass
class ProgressState:
    iters: numpy.ndarray = field(default_factory=lambda: numpy.array([], dtype=numpy.uint64))
    detector_errors: numpy.ndarray = field(default_factory=lambda: numpy.array([], dtype=numpy.float64))


@dataclass
class ReconsState:
    iter: IterState
    wavelength: float
    probe: ProbeState
    object: ObjectState
    scan: numpy.ndarray
    tilt: t.Optional[numpy.ndarray] = None
    progress: ProgressState = field(default_factory=ProgressState)

    def copy(self):
        return copy.deepcopy(self)
