"""Shared construction of device frameworks from the oracle's synthetic systems (tests only)."""
from __future__ import annotations

import numpy as np

import aleator_amd as aa
from oracle_lib import LJ, Framework


def device_framework(ctx, fw: Framework, lj: LJ, ewald=None) -> aa.Framework:
    return aa.Framework(ctx, fw.m, fw.x, fw.y, fw.z, fw.species, lj.eps, lj.sigma, lj.cutoff,
                        truncation=lj.truncation, mixing=lj.mixing, charge=fw.q, ewald=ewald)


def rel_err(gpu, ref, floor=1.0):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.abs(gpu - ref) / np.maximum(np.abs(ref), floor)


def sample_nodes(counts, n, seed):
    total = int(np.prod(counts))
    rng = np.random.default_rng(seed)
    nodes = rng.choice(total, size=min(n, total), replace=False)
    # always include the origin node and the last node (tile edges)
    return np.unique(np.concatenate([nodes, [0, total - 1]])).astype(np.int64)
