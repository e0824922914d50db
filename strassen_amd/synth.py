"""Seeded synthetic inputs for the block-sparse multiply workloads.

Element values are defined so that the host (numpy) and the device generator
kernel (``sa_bsm_generate``) produce bit-identical doubles:

* a counter-based hash: ``h = splitmix64(seed * 0xD1B54A32D192ED03 + 12*idx + t)``
  for ``t = 0..11``, where ``idx`` is the element's global column-major index
  ``r + c*nrow (+ l*nrow*ncol for a batch slice l)``;
* twelve 49-bit uniforms ``u_t = (h_t >> 15) * 2**-49`` summed in order, minus 6
  (Irwin-Hall approximation of N(0,1): mean 0, variance exactly 1).  Every
  partial sum is a multiple of 2**-49 below 2**4, so it is exact in FP64;
* one rounding: ``value = block_scale * z``.

Block scales (``exp(-0.5|I-J|)`` etc.) are computed once on the host and handed
to the device as a table, so no transcendental function is evaluated twice.

Geometry builders return block sizes, per-block scales (0 = absent block) and a
description; they correspond to the reference's benchmark shapes described in
main.cpp:10-68 (dense-derived uniform BSMs) and occ_ri_k_step.cpp:14-146.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1
SEED_MULT = 0xD1B54A32D192ED03
N_UNIFORM = 12
UNIFORM_BITS = 49


def _splitmix64(x: np.ndarray) -> np.ndarray:
    z = x + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def normal_like(seed: int, idx: np.ndarray) -> np.ndarray:
    """Irwin-Hall(12) - 6 values for global element indices ``idx`` (uint64)."""
    idx = np.asarray(idx, dtype=np.uint64)
    base = np.uint64((seed * SEED_MULT) & MASK64)
    acc = np.zeros(idx.shape, dtype=np.float64)
    scale = 2.0 ** -UNIFORM_BITS
    with np.errstate(over="ignore"):
        for t in range(N_UNIFORM):
            h = _splitmix64(base + idx * np.uint64(N_UNIFORM) + np.uint64(t))
            acc = acc + (h >> np.uint64(64 - UNIFORM_BITS)).astype(np.float64) * scale
    return acc - 6.0


def dense_from_scales(seed: int, rsz, csz, scale: np.ndarray, slice_index: int = 0) -> np.ndarray:
    """Dense column-major (Fortran-ordered) matrix whose block (I, J) holds
    ``scale[I, J] * normal_like(seed, global index)`` (zero where scale == 0)."""
    rsz = np.asarray(rsz, dtype=np.int64)
    csz = np.asarray(csz, dtype=np.int64)
    nr, nc = int(rsz.sum()), int(csz.sum())
    r = np.arange(nr, dtype=np.uint64)[:, None]
    c = np.arange(nc, dtype=np.uint64)[None, :]
    idx = r + c * np.uint64(nr) + np.uint64(slice_index) * np.uint64(nr * nc)
    z = normal_like(seed, idx)
    rb = np.repeat(np.arange(len(rsz)), rsz)
    cb = np.repeat(np.arange(len(csz)), csz)
    s = np.asarray(scale, dtype=np.float64)[rb][:, cb]
    out = s * z
    out[s == 0.0] = 0.0
    return np.asfortranarray(out)


def uniform_sizes(n: int, bs: int) -> np.ndarray:
    nb = (n + bs - 1) // bs
    s = np.full(nb, bs, dtype=np.int32)
    if nb and n % bs:
        s[-1] = n % bs
    return s


def band_scales(nb: int, band: int, decay: float) -> np.ndarray:
    """scale[I, J] = exp(-decay |I-J|) inside |I-J| <= band, else 0."""
    i = np.arange(nb)
    d = np.abs(i[:, None] - i[None, :])
    tab = np.array([math.exp(-decay * k) for k in range(nb + 1)])
    s = tab[d]
    s[d > band] = 0.0
    return s


@dataclass
class Workload:
    name: str
    rsz: np.ndarray            # row block sizes of A (and C)
    ksz: np.ndarray            # inner block sizes
    csz: np.ndarray            # column block sizes of B (and C)
    a_scale: np.ndarray        # per-block scale of A (0 = absent)
    b_scale: np.ndarray
    thresh: float
    uniform_bs: int = 0        # >0: uniform target block size (reference semantics)
    seeds: tuple = (1, 2)
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.rsz.sum())


def config1(n: int = 2048, bs: int = 64, band: int = 4, decay: float = 0.5, thresh: float = 1e-10) -> Workload:
    """n=2048, uniform 64x64 blocks, banded |I-J|<=4 (26.2 % fill),
    scale exp(-0.5|I-J|), threshold 1e-10, Strassen cutoff 512."""
    sz = uniform_sizes(n, bs)
    s = band_scales(len(sz), band, decay)
    return Workload("config1_band_fp64", sz, sz, sz, s, s, thresh, uniform_bs=bs,
                    extra={"band": band, "decay": decay, "cutoff": 512})


def config5(n: int = 32768, bs: int = 128, band: int = 95, decay: float = 0.05, thresh: float = 1e-10) -> Workload:
    """n=32768, uniform 128x128 blocks, band |I-J|<=95 (60.7 % fill),
    scale exp(-0.05|I-J|), Strassen depth <= 4 (cutoff 2048)."""
    sz = uniform_sizes(n, bs)
    s = band_scales(len(sz), band, decay)
    return Workload("config5_hybrid", sz, sz, sz, s, s, thresh, uniform_bs=bs,
                    extra={"band": band, "decay": decay, "cutoff": 2048})


SHELL_SIZES = np.array([5, 9, 14, 18, 24, 30, 40], dtype=np.int32)


def atom_sizes(n_target: int, rng: np.random.Generator, table=SHELL_SIZES) -> np.ndarray:
    """Draw per-atom block sizes from ``table`` until the total reaches
    ``n_target``; the last size is trimmed so the sum is exact."""
    sizes = []
    tot = 0
    while tot < n_target:
        s = int(rng.choice(table))
        s = min(s, n_target - tot)
        sizes.append(s)
        tot += s
    if len(sizes) > 1 and sizes[-1] < int(table.min()):
        sizes[-2] += sizes[-1]
        sizes.pop()
    return np.asarray(sizes, dtype=np.int32)


def config2(n: int = 12000, fill: float = 0.10, decay: float = 0.3, thresh: float = 1e-10,
            seed: int = 3) -> Workload:
    """Uneven atom-shaped blocks (5-40, seeded shell-size table), n about 12000,
    ~10 % block fill by distance cutoff over random 3D atom positions, block
    scale exp(-0.3 d_IJ)."""
    rng = np.random.default_rng(seed)
    sz = atom_sizes(n, rng)
    na = len(sz)
    box = (na / 0.05) ** (1.0 / 3.0)  # number density 0.05 atoms / unit^3
    pos = rng.uniform(0.0, box, size=(na, 3))
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    rc = float(np.quantile(d.ravel(), fill))
    s = np.exp(-decay * d)
    s[d > rc] = 0.0
    return Workload("config2_uneven_atoms", sz, sz, sz, s, s, thresh,
                    extra={"natoms": na, "rc": rc, "box": box, "decay": decay, "fill": fill})


def small_uneven(n: int = 600, fill: float = 0.3, seed: int = 7, thresh: float = 1e-10) -> Workload:
    w = config2(n=n, fill=fill, seed=seed, thresh=thresh)
    w.name = "small_uneven"
    return w
