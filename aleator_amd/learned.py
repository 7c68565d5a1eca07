"""Learned insertion density for GCMC (no reference counterpart).

The method's own energy grids train a Fourier-feature MLP (aleator_gpu_mlp_*, BF16 MFMA on the
device) to predict a probe energy from a fractional position; its predictions at the centres of
a bin table over the primitive cell give bin weights exp(-(U - U_min) / T_p), which
Gcmc.set_proposal turns into the insertion density and corrects acceptance for
(aleator_gpu_gcmc_set_proposal).  The exact energy decides every acceptance, so the density
only changes how fast a chain samples, never what it samples.

Training set: n nodes of the grid drawn with replacement (node = floor(u N) with u from
Philox4x32Rng(seed, stream) at calls 0..n-1), each with its fractional position (i/n0, j/n1,
k/n2) and a combined energy sum_t c_t U_t (for a rigid CO2: U_C + 2 U_O, the molecule with its
three sites collapsed onto the centre; the Coulomb potential drops out because the molecule is
neutral).  Targets are clipped at e_clip (the repulsive cores carry no information the density
needs beyond "very unlikely") and standardised.
"""
from __future__ import annotations

import numpy as np

from . import Context, DeviceBuffer, Mlp, philox_uniforms


def node_fractions(n_points, nodes) -> np.ndarray:
    """Fractional positions of flat grid nodes ((iz*ny + iy)*nx + ix), double [n][3]."""
    nx, ny, nz = (int(v) for v in n_points)
    nodes = np.asarray(nodes, dtype=np.int64)
    ix = nodes % nx
    iy = (nodes // nx) % ny
    iz = nodes // (nx * ny)
    return np.stack([ix / nx, iy / ny, iz / nz], axis=1)


def sample_grid(tables, coefs, n_points, n_samples: int, seed: int = 4, stream: int = 0,
                low_fraction: float = 0.0, low_cut: float = 0.0):
    """(fractional positions [n][3], energies [n]) of n_samples grid nodes drawn with
    replacement; tables [t][nz][ny][nx], energy = sum_t coefs[t] * tables[t].

    With low_fraction > 0 that share of the samples is drawn from the nodes whose energy is below
    low_cut (Philox stream + 1) and the rest from all nodes: the regions an insertion density
    has to resolve are a small part of a dense framework's volume."""
    total = int(np.prod(n_points))
    energy_all = np.zeros(total)
    for c, t in zip(coefs, tables):
        energy_all += c * np.asarray(t, dtype=np.float64).reshape(-1)
    n_low = int(round(low_fraction * n_samples))
    low = np.flatnonzero(energy_all < low_cut) if n_low else np.zeros(0, dtype=np.int64)
    if len(low) == 0:
        n_low = 0
    u = philox_uniforms(seed, stream, 0, n_samples - n_low)
    nodes = np.minimum((u * total).astype(np.int64), total - 1)
    if n_low:
        ul = philox_uniforms(seed, stream + 1, 0, n_low)
        nodes = np.concatenate([nodes, low[np.minimum((ul * len(low)).astype(np.int64), len(low) - 1)]])
    return node_fractions(n_points, nodes), energy_all[nodes]


class LearnedDensity:
    """An MLP fitted to a framework's probe energy and the insertion density it defines."""

    def __init__(self, ctx: Context, max_rows: int = 8192, seed: int = 0):
        self.net = Mlp(ctx, max_rows=max_rows, seed=seed)
        self.mean = 0.0
        self.scale = 1.0

    def fit(self, frac, energy, e_clip: float = 2000.0, n_iters: int = 2000, batch: int = 4096, lr: float = 1e-3,
            seed: int = 5, stream: int = 0) -> np.ndarray:
        """Trains on (frac, min(energy, e_clip)) standardised; returns the per-iteration loss."""
        target = np.minimum(np.asarray(energy, dtype=np.float64), e_clip)
        self.mean = float(target.mean())
        self.scale = float(target.std()) or 1.0
        d_f = DeviceBuffer.from_array(np.ascontiguousarray(frac, dtype=np.float64))
        d_t = DeviceBuffer.from_array((target - self.mean) / self.scale)
        try:
            return self.net.train_device(int(d_f), int(d_t), len(target), batch, n_iters, lr, seed=seed,
                                         stream=stream)
        finally:
            d_f.free()
            d_t.free()

    def energy(self, frac) -> np.ndarray:
        """Predicted energy (K) at fractional positions."""
        return self.mean + self.scale * self.net.predict(frac).astype(np.float64)

    def bin_centres(self, nb) -> np.ndarray:
        i, j, k = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in nb], indexing="ij")
        return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)  # row-major, axis 2 fastest

    def bin_weights(self, nb, temperature: float) -> np.ndarray:
        """exp(-(U - U_min) / T) at the bin centres, in the row-major order set_proposal takes."""
        u = self.energy(self.bin_centres(nb))
        return np.exp(-(u - u.min()) / temperature)
