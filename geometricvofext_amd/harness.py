"""Time control and test-case set-up of the reference's advection benchmark
(``applications/test/plicVofAdvectionFoam``).

``RunTime`` restates the pieces of OpenFOAM's ``Time`` that the harness drives
(``setDeltaT.H:34-53`` + ``Time::adjustDeltaT`` for ``writeControl
adjustableRunTime`` + ``Time::operator++`` + ``Time::run``), with the controls
of ``tutorials/test/plicVofAdvectionFoam/system/controlDict``.  Function-object
time-step adjustment (sampling surfaces) is out of scope.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

SMALL = 1e-15


@dataclass
class RunTime:
    end_time: float = 3.0
    delta_t: float = 0.001
    write_interval: float = 0.25
    start_time: float = 0.0
    max_co: float = 0.5
    max_alpha_co: float = 0.5
    max_delta_t: float = 0.2
    adjust_time_step: bool = True

    def __post_init__(self):
        self.value = self.start_time
        self.time_index = 0
        self.write_time_index = 0
        self.write_time = False

    def run(self) -> bool:
        # Time::run: value() < endTime - 0.5*deltaT
        return self.value < (self.end_time - 0.5 * self.delta_t)

    def set_initial_delta_t(self, co_num: float) -> None:
        # setInitialDeltaT.H
        if self.adjust_time_step and self.time_index == 0 and co_num > SMALL:
            self.delta_t = min(self.max_co * self.delta_t / co_num, min(self.delta_t, self.max_delta_t))

    def _adjust_delta_t(self) -> None:
        # Time::adjustDeltaT for adjustableRunTime writes
        time_to_next_write = max(0.0, (self.write_time_index + 1) * self.write_interval
                                 - (self.value - self.start_time))
        n_steps = time_to_next_write / self.delta_t
        if n_steps < 2147483647.0:
            n_to_write = int(max(1.0, _round_half_away(n_steps)))
            new_dt = time_to_next_write / n_to_write
            if new_dt >= self.delta_t:
                self.delta_t = min(new_dt, 2.0 * self.delta_t)
            else:
                self.delta_t = max(new_dt, 0.2 * self.delta_t)

    def set_delta_t(self, co_num: float, alpha_co_num: float) -> None:
        # setDeltaT.H:34-53
        if not self.adjust_time_step:
            return
        max_fact = min(self.max_co / (co_num + SMALL), self.max_alpha_co / (alpha_co_num + SMALL))
        fact = min(min(max_fact, 1.0 + 0.1 * max_fact), 1.2)
        self.delta_t = min(fact * self.delta_t, self.max_delta_t)
        self._adjust_delta_t()

    def increment(self) -> None:
        # Time::operator++
        self.value = self.value + self.delta_t
        self.time_index += 1
        if abs(self.value) < 10 * SMALL * self.delta_t:
            self.value = 0.0
        self.write_time = False
        write_index = int(((self.value - self.start_time) + 0.5 * self.delta_t) / self.write_interval)
        if write_index > self.write_time_index:
            self.write_time = True
            self.write_time_index = write_index


def _round_half_away(x: float) -> float:
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def u_factor(t: float, dt: float, period: float) -> float:
    """updateU.H:117-127: time factor of the deformation field (post-increment t)."""
    if period <= 0.0:
        return 1.0
    return 0.5 * (math.cos(2.0 * math.pi * t / period) + math.cos(2.0 * math.pi * (t + dt) / period))


def shape_error(alpha, alpha_exact, V) -> float:
    """E_s = sum|alpha - alpha_exact| V / sum alpha_exact V (updateErrors.H:31)."""
    import numpy as np
    return float(np.sum(np.abs(alpha - alpha_exact) * V) / np.sum(alpha_exact * V))


def volume_error(alpha, V, V0) -> float:
    """E_v = (sum alpha V - V0) / V0 (updateErrors.H)."""
    import numpy as np
    return float((np.sum(alpha * V) - V0) / V0)
