"""EpsilonScalingHandler (include/SemiDiscreteOT/solvers/EpsilonScalingHandler.h,
src/solvers/EpsilonScalingHandler.cc): the epsilon schedule of a solve and its
distribution over the levels of a multilevel run.  Host-side bookkeeping; the
solves themselves run on the GPU."""
from __future__ import annotations

import sys


def _warn(msg: str):
    print(msg, file=sys.stderr)


class EpsilonScalingHandler:
    def __init__(self, initial_epsilon: float, scaling_factor: float, num_steps: int):
        """EpsilonScalingHandler.cc:6-28: a factor <= 1 falls back to 2 and
        zero steps to 1, with a warning."""
        self.initial_epsilon = float(initial_epsilon)
        self.scaling_factor = float(scaling_factor)
        self.num_steps = int(num_steps)
        if self.scaling_factor <= 1.0:
            _warn("Warning: Epsilon scaling factor should be > 1.0. Using default value of 2.0.")
            self.scaling_factor = 2.0
        if self.num_steps == 0:
            _warn("Warning: Number of epsilon scaling steps cannot be 0. Using default value of 1.")
            self.num_steps = 1
        self.epsilon_distribution: list[list[float]] = []

    def generate_epsilon_sequence(self) -> list[float]:
        """EpsilonScalingHandler.cc:30-42: largest to smallest,
        initial_epsilon * factor^(steps - 1 - i)."""
        n = self.num_steps
        return [self.initial_epsilon * self.scaling_factor ** float(n - 1 - i) for i in range(n)]

    def distribute_epsilon_values(self, epsilon_sequence, num_levels: int) -> list[list[float]]:
        """EpsilonScalingHandler.cc:44-90: one value per level from the
        coarsest (the smallest repeated on the remaining levels) when there
        are no more values than levels; otherwise ceil(remaining values /
        remaining levels) values per level, coarsest first."""
        dist: list[list[float]] = [[] for _ in range(num_levels)]
        if num_levels == 0:
            _warn("Warning: No levels provided for epsilon distribution.")
            return dist
        seq = list(epsilon_sequence)
        if len(seq) <= num_levels:
            for level in range(num_levels):
                dist[level].append(seq[level] if level < len(seq) else seq[-1])
        else:
            remaining_eps, remaining_levels, k = len(seq), num_levels, 0
            for level in range(num_levels):
                n_here = (remaining_eps + remaining_levels - 1) // remaining_levels
                for _ in range(n_here):
                    if k < len(seq):
                        dist[level].append(seq[k])
                        k += 1
                remaining_eps -= n_here
                remaining_levels -= 1
        return dist

    def compute_epsilon_distribution(self, num_levels: int) -> list[list[float]]:
        """EpsilonScalingHandler.cc:92-100."""
        self.epsilon_distribution = self.distribute_epsilon_values(
            self.generate_epsilon_sequence(), num_levels)
        return self.epsilon_distribution

    def get_epsilon_values_for_level(self, level_index: int) -> list[float]:
        """EpsilonScalingHandler.cc:102-114: empty (with a warning) for an
        index beyond the distribution."""
        if level_index >= len(self.epsilon_distribution):
            _warn(f"Warning: Requested epsilon values for invalid level index: {level_index}")
            return []
        return self.epsilon_distribution[level_index]

    def print_epsilon_distribution(self, out=sys.stdout):
        """EpsilonScalingHandler.cc:116-159."""
        if not self.epsilon_distribution:
            print("Epsilon distribution has not been computed yet.", file=out)
            return
        n = len(self.epsilon_distribution)
        print("\n----------------------------------------", file=out)
        print("Epsilon Scaling Distribution:", file=out)
        print("----------------------------------------", file=out)
        print(f"Initial epsilon: {self.initial_epsilon:g}", file=out)
        print(f"Scaling factor: {self.scaling_factor:g}", file=out)
        print(f"Number of steps: {self.num_steps}", file=out)
        print(f"Number of levels: {n}", file=out)
        print("----------------------------------------", file=out)
        for level, eps in enumerate(self.epsilon_distribution):
            kind = "coarsest" if level == 0 else ("finest" if level == n - 1 else "intermediate")
            vals = ", ".join(f"{e:.4e}" for e in eps)
            print(f"Level {n - 1 - level} ({kind}): {len(eps)} epsilon value(s): {vals}", file=out)
        print("----------------------------------------", file=out)
