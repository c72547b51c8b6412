"""Epsilon schedules of a regularized solve and their split over the levels of
a multilevel run (the schedule SemiDiscreteOT::run_sot and the multilevel
drivers take from EpsilonScalingHandler, src/solvers/EpsilonScalingHandler.cc:
6-100).  Host-side arithmetic only; the solves run on the GPU.

Two rules fix the numbers and are kept for parity:
  * a schedule of n stages ends at the requested epsilon and grows by the
    factor towards the first stage: stage i solves at eps * factor^(n-1-i);
    a factor <= 1 means 2 and 0 stages mean 1 (EpsilonScalingHandler.cc:18-27);
  * over L levels (coarsest first) a schedule of at most L values gives one
    value per level, the last value repeated on the levels beyond it; a longer
    one gives each level in turn ceil(values left / levels left) values
    (EpsilonScalingHandler.cc:44-90).
"""
from __future__ import annotations


def normalized_scaling(factor: float, steps: int) -> tuple[float, int]:
    """(factor, steps) after the fallbacks: factor <= 1 -> 2, steps 0 -> 1."""
    f = float(factor)
    n = int(steps)
    return (f if f > 1.0 else 2.0), (n if n != 0 else 1)


def epsilon_schedule(epsilon: float, factor: float, steps: int) -> list[float]:
    """Stage values from the largest to the requested epsilon."""
    f, n = normalized_scaling(factor, steps)
    eps = float(epsilon)
    return [eps * f ** float(n - 1 - i) for i in range(n)]


def split_schedule(schedule, n_levels: int) -> list[list[float]]:
    """The stages each level solves, coarsest level first."""
    vals = [float(v) for v in schedule]
    if n_levels <= 0:
        return []
    if not vals:
        return [[] for _ in range(n_levels)]
    if len(vals) <= n_levels:
        return [[vals[min(k, len(vals) - 1)]] for k in range(n_levels)]
    out, pos = [], 0
    for k in range(n_levels):
        levels_left = n_levels - k
        take = -(-(len(vals) - pos) // levels_left)
        out.append(vals[pos:pos + take])
        pos += take
    return out


def level_schedules(epsilon: float, factor: float, steps: int, n_levels: int) -> list[list[float]]:
    """epsilon_schedule split over n_levels levels."""
    return split_schedule(epsilon_schedule(epsilon, factor, steps), n_levels)


def stages_of_level(schedules: list[list[float]], level: int, fallback: float) -> list[float]:
    """The stages of one level, or [fallback] where the split gives none."""
    if 0 <= level < len(schedules) and schedules[level]:
        return list(schedules[level])
    return [float(fallback)]
