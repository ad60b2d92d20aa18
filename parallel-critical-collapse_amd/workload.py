"""Workload definitions: the five benchmark shapes and their seeded inputs.

The Newton vector x (NewtonSolver.cpp:273-274) of every problem is

    x = x_anchor(N) + delta(seed)

x_anchor(N) is the reference's shipped near-solution data/simulation_4D_<N>.json
packed with ICG.cpp:678-711 (stored as a fixture, tests/golden/anchor_4D_<N>.json).
delta is a seeded perturbation with the requested distribution: real periodic
series P(tau) = sum_k a_k cos(2 pi k j/N) + b_k sin(2 pi k j/N), a_k, b_k drawn
N(0,1) * 0.1/k^2, on odd harmonics k = 1, 3, ... < N/4 for Up and psic and on
even harmonics k = 2, 4, ... < N/4 for fc (fc keeps only even harmonics in the
packing, ICG.cpp:705-710).  Under forwardFFT (exp(+2 pi i jk/N)/N) such a series
has coefficient (a_k + i b_k)/2 at mode k, so delta is written directly in the
packed layout.  Slot N/2+2 (Re of fc's mode 2) carries Delta instead
(NewtonSolver.cpp:274), so a_2 of fc is dropped there.

Random numbers: SplitMix64 -> Box-Muller (cos branch, two uniforms per draw),
draw order Up, psic, fc, k ascending, a_k before b_k.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field, replace

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SEED_BASE = 20261203


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


@dataclass(frozen=True)
class Shape:
    """One solver configuration (the SimulationConfig keys of the path)."""
    Ntau: int
    X: int                       # NLeft + NRight
    Dim: float = 4.0
    XLeft: float = 1e-3
    XMid: float = 0.5
    XRight: float = 1.0 - 1e-3
    Delta: float = 3.44
    SchemeIRK: int = 2
    PrecisionIRK: float = 1e-15
    MaxIterIRK: int = 100
    EpsNewton: float = 1e-8
    PrecisionNewton: float = 1e-10
    SlowError: float = 1e-3
    MaxIterNewton: int = 100
    seed: int = SEED_BASE
    scale: float = 1.0           # perturbation scale

    @property
    def NLeft(self) -> int:
        return self.X // 2

    @property
    def NRight(self) -> int:
        return self.X - self.X // 2

    @property
    def n(self) -> int:
        return 3 * self.Ntau // 4


# The five shapes of the benchmark plan.  C4 is a list of 24 problems.
def config(index: int) -> Shape:
    if index == 1:
        return Shape(Ntau=128, X=1000, seed=SEED_BASE + 1)
    if index == 2:
        return Shape(Ntau=512, X=6000, seed=SEED_BASE + 2)
    if index == 3:
        # scale 1 makes the FP64 oracle's right-side IRK iteration fail near
        # x = 0.7; 10^-1 is the first power of ten it survives (k = 5.1)
        return Shape(Ntau=1024, X=8000, seed=SEED_BASE + 3, scale=0.1)
    if index == 5:
        # as for index 3: the oracle's FP64 shoot fails at scale 1, survives at 10^-1
        return Shape(Ntau=2048, X=12000, seed=SEED_BASE + 5, scale=0.1)
    raise ValueError(index)


# Perturbation scale of each batched problem: the largest power of ten <= 1
# at which the oracle's FP64 shoot of the seeded input survives.
CONFIG4_SCALE = (0.1, 0.1, 0.1, 0.1, 1.0, 1.0, 1.0, 1.0, 1.0, 0.1, 1.0, 0.1,
                 0.1, 1.0, 1.0, 0.1, 1.0, 0.1, 0.1, 1.0, 0.1, 1.0, 0.1, 0.1)


def config4() -> list[Shape]:
    """24 uneven problems: Ntau and X drawn by seed from {128, 256, 512} and
    {2000, 4000, 6000}, Dim = 3.96 + 0.01 (i mod 15), anchored on the D=4 data.

    Two measured feasibility limits shape this set: at Ntau=512 the
    reference's fixed-point IRK diverges on the 2000-point grid even for the
    unperturbed D=4 data (so X is raised to 4000 there), and with D=4 data
    as the starting point every shoot with D <= 3.93 or D >= 4.6 that was
    tried failed at every perturbation scale, as did most with D <= 3.9."""
    rng = SplitMix64(SEED_BASE + 4)
    out = []
    for i in range(24):
        Ntau = (128, 256, 512)[rng.next_u64() % 3]
        X = (2000, 4000, 6000)[rng.next_u64() % 3]
        if Ntau == 512 and X < 4000:
            X = 4000
        out.append(Shape(Ntau=Ntau, X=X, Dim=round(3.96 + 0.01 * (i % 15), 10),
                         seed=SEED_BASE + 4000 + i, scale=CONFIG4_SCALE[i]))
    return out


def anchor(Ntau: int) -> tuple[np.ndarray, float]:
    with open(os.path.join(GOLDEN, f"anchor_4D_{Ntau}.json")) as f:
        d = json.load(f)
    return np.asarray(d["x"], np.float64), float(d["Delta"])


def perturbation(Ntau: int, seed: int, scale: float = 1.0) -> np.ndarray:
    n = 3 * Ntau // 4
    rng = SplitMix64(seed)
    d = np.zeros(n)
    # Up: modes k = 2j+1 at slots (2j, 2j+1); psic the same at offset N/4.
    for off in (0, Ntau // 4):
        for k in range(1, Ntau // 4, 2):
            a = rng.normal() * 0.1 / k ** 2
            b = rng.normal() * 0.1 / k ** 2
            d[off + k - 1] += scale * a / 2
            d[off + k] += scale * b / 2
    # fc: modes k = 2j at slots N/2 + (2j, 2j+1).
    for k in range(2, Ntau // 4, 2):
        a = rng.normal() * 0.1 / k ** 2
        b = rng.normal() * 0.1 / k ** 2
        d[Ntau // 2 + k] += scale * a / 2
        d[Ntau // 2 + k + 1] += scale * b / 2
    return d


def initial_x(shape: Shape) -> np.ndarray:
    """Seeded anchored Newton vector with Delta in slot N/2+2."""
    xa, _ = anchor(shape.Ntau)
    x = xa + perturbation(shape.Ntau, shape.seed, shape.scale)
    x[shape.Ntau // 2 + 2] = shape.Delta
    return x


def grid(shape: Shape) -> np.ndarray:
    """NewtonSolver.cpp:477-519: uniform in z = ln x - ln(1-x)."""
    iM, iR = shape.NLeft, shape.X
    g = np.zeros(shape.X + 1)
    g[0], g[iM], g[iR] = shape.XLeft, shape.XMid, shape.XRight
    zl = math.log(shape.XLeft) - math.log(1.0 - shape.XLeft)
    zm = math.log(shape.XMid) - math.log(1.0 - shape.XMid)
    zr = math.log(shape.XRight) - math.log(1.0 - shape.XRight)
    dzl = (zm - zl) / shape.NLeft
    dzr = (zr - zm) / shape.NRight
    for j in range(1, iM):
        e = zl + j * dzl
        g[j] = math.exp(e) / (1.0 + math.exp(e))
    for j in range(iM + 1, iR):
        e = zm + (j - iM) * dzr
        g[j] = math.exp(e) / (1.0 + math.exp(e))
    return g
