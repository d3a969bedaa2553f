"""Seeded synthetic workloads of the benchmark configurations.

Each function returns databases built by the native generators
(include/gspan_amd.h: gm_gen_random / gm_gen_planted / gm_gen_molecule)
with the sizes, label alphabets and support thresholds of the five
benchmark configurations:

  config 0  1,000 random connected graphs (V ~ U[10,30], E = round(1.2 V),
            5 vertex / 3 edge labels), seed 0, min support 10 %
  config 1  10,000 planted-pattern transactions (Poisson(30) edges, 50 seed
            patterns of max(1, Poisson(4)) edges, 5/5 labels, merge 0.5),
            seed 2, min support 5 %
  config 2  the same generator at 100,000 transactions, seed 3, 1 %
  config 3  256 jobs, |D| log-uniform in [200, 20000] drawn from the
            config-0 generator with V ~ U[10,60], min support U[2 %, 20 %],
            seed 1
  config 4  50,000 molecule-like graphs (V ~ Poisson(25), degree <= 4,
            1-3 rings of 5-6), seed 4
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

from . import GraphDB


def abs_support(frac: float, n_graphs: int) -> int:
    """Absolute support from a fraction: ceil(frac * |D|), at least 1."""
    return max(1, int(math.ceil(frac * n_graphs - 1e-9)))


def config0(seed: int = 0, n_graphs: int = 1000):
    return GraphDB.random(seed, n_graphs, 10, 30, 1.2, 5, 3), abs_support(0.10, n_graphs)


def config1(seed: int = 2, n_graphs: int = 10000):
    return GraphDB.planted(seed, n_graphs, 30.0, 50, 4.0, 5, 5, 0.5), abs_support(0.05, n_graphs)


def config2(seed: int = 3, n_graphs: int = 100000):
    return GraphDB.planted(seed, n_graphs, 30.0, 50, 4.0, 5, 5, 0.5), abs_support(0.01, n_graphs)


class _SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014): the same stream the native
    generators use, so job lists are identical on every machine."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def unit(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


@dataclass(frozen=True)
class Job:
    index: int
    seed: int
    n_graphs: int
    min_support_frac: float
    minsup: int

    def database(self) -> GraphDB:
        return GraphDB.random(self.seed, self.n_graphs, 10, 60, 1.2, 5, 3)


def config3_jobs(n_jobs: int = 256, seed: int = 1, min_graphs: int = 200, max_graphs: int = 20000) -> List[Job]:
    """The 256-job mix: |D| log-uniform in [min_graphs, max_graphs], min
    support uniform in [2 %, 20 %]."""
    r = _SplitMix64(seed)
    jobs = []
    lo, hi = math.log(min_graphs), math.log(max_graphs)
    for j in range(n_jobs):
        n = int(round(math.exp(lo + (hi - lo) * r.unit())))
        frac = 0.02 + 0.18 * r.unit()
        jseed = r.next()
        jobs.append(Job(j, jseed, n, frac, abs_support(frac, n)))
    return jobs


def config4(seed: int = 4, n_graphs: int = 50000) -> GraphDB:
    return GraphDB.molecule(seed, n_graphs, 25.0, 4)
