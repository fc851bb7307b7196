"""The t0 lattice on which the fused gather evaluates the traveltime network.

Nodes are output-sample indices n_0 = 0, n_{j+1} = min(n_j + min(max((n_j g)
>> 8, K), K_max), nt - 1), up to the last sample nt - 1: spacing K at shallow
times, growing by g / 256 of the node's own index, capped at K_max.  g = 0
gives the uniform lattice K, 2K, ... (the last interval shorter).  Each output
sample takes the network's outputs from the cubic Hermite polynomial of its
segment [n_j, n_{j+1}] (values at the two nodes, slopes from differences
over nodes j - 1 .. j + 2).  The same integer law runs in the gather
(lattice_next in pstm_hip.hip) and here, for the held-out check and the
tests.

Why a growing spacing: the learned factor of a leg changes fastest where
t0 / 2 is comparable to the leg's horizontal time d / v, which for the short
legs that dominate the interpolation error is at shallow times; deeper, it
is smooth on the scale of tens of samples.  A lattice dense at shallow times
and sparse below reaches the interpolation error of a uniform lattice with
far fewer network evaluations per (pillar, trace) pair.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Lattice:
    spacing: int = 16       # K: smallest node spacing (output samples), 4 .. 512
    growth: int = 0         # g: spacing growth in 1/256 of the node index, 0 .. 1024
    max_spacing: int = 0    # K_max: largest spacing, K .. 1024 (0: K)

    def __post_init__(self):
        if not 4 <= self.spacing <= 512:
            raise ValueError("node spacing must be in [4, 512]")
        if not 0 <= self.growth <= 1024:
            raise ValueError("node growth must be in [0, 1024]")
        if self.max_spacing and not self.spacing <= self.max_spacing <= 1024:
            raise ValueError("largest node spacing must be 0 or in [spacing, 1024]")

    @property
    def kmax(self) -> int:
        return self.max_spacing if self.max_spacing > 0 else self.spacing

    def nodes(self, nt: int) -> np.ndarray:
        """Node sample indices over an axis of nt samples (int64, ascending,
        first 0, last nt - 1)."""
        if nt < 2:
            raise ValueError("the lattice needs at least 2 output samples")
        pos = [0]
        while pos[-1] < nt - 1:
            n = pos[-1]
            step = min(max((n * self.growth) >> 8, self.spacing), self.kmax)
            pos.append(min(n + step, nt - 1))
        return np.asarray(pos, dtype=np.int64)


def as_lattice(node_spacing) -> Lattice:
    """A Lattice from a Lattice or an int (uniform spacing)."""
    if isinstance(node_spacing, Lattice):
        return node_spacing
    return Lattice(spacing=int(node_spacing))


# a growing lattice for general use: 8 samples at shallow times, growing by
# 77/256 (~0.3) of the node index, at most 48 samples (40 nodes over 1500
# samples where the uniform 16-sample lattice has 95)
GEOMETRIC = Lattice(spacing=8, growth=77, max_spacing=48)

# the benchmark's lattice (and the full-size VTI test's): growing by 100/256
# of the node index, at most 64 samples (31 nodes over 1500 samples); on the
# full-size VTI job, with a quantisation-aware-trained network, its image
# error against the exact traveltime is at the network's own floor (denser
# lattices gain nothing), while the 23-node (8, 128, 96) lattice adds to it
SPARSE = Lattice(spacing=8, growth=100, max_spacing=64)
