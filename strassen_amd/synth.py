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


def band_workload(n: int = 2048, bs: int = 64, band: int = 4, decay: float = 0.5, thresh: float = 1e-10) -> Workload:
    """n=2048, uniform 64x64 blocks, banded |I-J|<=4 (26.2 % fill),
    scale exp(-0.5|I-J|), threshold 1e-10, Strassen cutoff 512."""
    sz = uniform_sizes(n, bs)
    s = band_scales(len(sz), band, decay)
    return Workload(f"band_n{n}_bs{bs}_w{band}", sz, sz, sz, s, s, thresh, uniform_bs=bs,
                    extra={"band": band, "decay": decay, "cutoff": 512})


def hybrid_workload(n: int = 32768, bs: int = 128, band: int = 95, decay: float = 0.05, thresh: float = 1e-10) -> Workload:
    """n=32768, uniform 128x128 blocks, band |I-J|<=95 (60.7 % fill),
    scale exp(-0.05|I-J|), Strassen depth <= 4 (cutoff 2048)."""
    sz = uniform_sizes(n, bs)
    s = band_scales(len(sz), band, decay)
    return Workload(f"hybrid_n{n}_bs{bs}_band{band}", sz, sz, sz, s, s, thresh, uniform_bs=bs,
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


def atoms_workload(n: int = 12000, fill: float = 0.10, decay: float = 0.3, thresh: float = 1e-10,
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
    return Workload(f"uneven_atoms_n{n}_fill{fill:g}", sz, sz, sz, s, s, thresh,
                    extra={"natoms": na, "rc": rc, "box": box, "decay": decay, "fill": fill})


def small_uneven(n: int = 600, fill: float = 0.3, seed: int = 7, thresh: float = 1e-10) -> Workload:
    w = atoms_workload(n=n, fill=fill, seed=seed, thresh=thresh)
    w.name = "small_uneven"
    return w


# ------------------------------------------------------------------ RI-K --
BF_SIZES = np.array([9, 13, 14, 18], dtype=np.int32)
AUX_SIZES = np.array([30, 36, 40, 44, 50], dtype=np.int32)


def _exact_sizes(rng, natoms, total, table):
    s = rng.choice(table, size=natoms).astype(np.int64)
    # scale to the exact total, keeping every size >= 1
    s = np.maximum(1, np.round(s * (total / s.sum()))).astype(np.int64)
    d = total - int(s.sum())
    i = 0
    while d != 0:
        j = i % natoms
        if d > 0:
            s[j] += 1; d -= 1
        elif s[j] > 1:
            s[j] -= 1; d += 1
        i += 1
    return s.astype(np.int32)


def _morton_order(pos, box):
    q = np.clip((pos / box * 1023).astype(np.int64), 0, 1023)
    code = np.zeros(len(pos), dtype=np.int64)
    for b in range(10):
        for d in range(3):
            code |= ((q[:, d] >> b) & 1) << (3 * b + d)
    return np.argsort(code, kind="stable")


def _groups(sizes, target):
    """Consecutive items grouped until each group holds >= target functions."""
    out, cur, acc = [], [], 0
    for i, s in enumerate(sizes):
        cur.append(i); acc += int(s)
        if acc >= target:
            out.append(cur); cur, acc = [], 0
    if cur:
        if out and acc < target // 2:
            out[-1].extend(cur)
        else:
            out.append(cur)
    return out


@dataclass
class RIKWorkload:
    name: str
    natoms: int
    n_bf: int
    n_aux: int
    n_occ: int
    p_sizes: np.ndarray       # aux functions per P-group (row blocks of B, X)
    mu_sizes: np.ndarray      # basis functions per mu-group
    i_chunks: np.ndarray      # occupied-index chunks
    b_scale: np.ndarray       # (n_pgroups, n_mugroups * n_ic) block scales of B, column block ic + n_ic*mu
    aux_pos: np.ndarray       # (n_aux, 3) positions of the aux functions (P-group order)
    alpha: float
    ridge: float
    thresh: float
    extra: dict = field(default_factory=dict)

    @property
    def b_col_sizes(self):
        return np.array([int(m) * int(i) for m in self.mu_sizes for i in self.i_chunks], dtype=np.int32)

    def metric(self):
        """M_PQ = exp(-alpha |r_P - r_Q|^2) + ridge delta_PQ (SPD), column-major."""
        r = self.aux_pos
        sq = (r * r).sum(1)
        d2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * (r @ r.T), 0.0)
        M = np.exp(-self.alpha * d2)
        M[np.diag_indices_from(M)] += self.ridge
        return np.asfortranarray(M)


def rik_config(natoms=300, n_bf=4000, n_aux=12000, n_occ=200, occ_chunk=50, p_group=160, mu_group=100,
               fill=0.35, decay=0.3, alpha=0.5, ridge=0.05, jitter=0.35, thresh=1e-10, seed=4,
               name="ri_k_step_300_atoms") -> RIKWorkload:
    """RI-K step: 300 random atoms, n_bf = 4000, n_aux = 12000, n_occ = 200.
    Atoms are ordered along a Morton curve and grouped into P-groups of about
    p_group aux functions and mu-groups of about mu_group basis functions; B
    block (P-group, mu-group) is present when the closest atom pair of the two
    groups lies within the radius that gives the requested block fill, with
    scale exp(-decay d_min)."""
    rng = np.random.default_rng(seed)
    box = (natoms / 0.05) ** (1.0 / 3.0)
    pos = rng.uniform(0.0, box, size=(natoms, 3))
    pos = pos[_morton_order(pos, box)]
    bf = _exact_sizes(rng, natoms, n_bf, BF_SIZES)
    aux = _exact_sizes(rng, natoms, n_aux, AUX_SIZES)
    pg = _groups(aux, p_group)
    mg = _groups(bf, mu_group)
    p_sizes = np.array([int(aux[g].sum()) for g in pg], dtype=np.int32)
    mu_sizes = np.array([int(bf[g].sum()) for g in mg], dtype=np.int32)
    d_atoms = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    dmin = np.array([[d_atoms[np.ix_(a, b)].min() for b in mg] for a in pg])
    rc = float(np.quantile(dmin.ravel(), fill))
    s = np.exp(-decay * dmin)
    s[dmin > rc] = 0.0
    n_ic = -(-n_occ // occ_chunk)
    i_chunks = np.array([min(occ_chunk, n_occ - i * occ_chunk) for i in range(n_ic)], dtype=np.int32)
    b_scale = np.repeat(s, n_ic, axis=1)  # column block ic + n_ic*mu
    aux_pos = np.repeat(pos, aux, axis=0) + rng.normal(0.0, jitter, size=(n_aux, 3))
    return RIKWorkload(name, natoms, n_bf, n_aux, n_occ, p_sizes, mu_sizes, i_chunks, b_scale, aux_pos, alpha,
                       ridge, thresh, extra={"rc": rc, "box": box, "fill": fill, "decay": decay})


def rik_small(seed=5):
    """A scaled-down RI-K geometry for parity tests (all sizes small)."""
    return rik_config(natoms=24, n_bf=240, n_aux=600, n_occ=24, occ_chunk=8, p_group=60, mu_group=40, fill=0.5,
                      seed=seed, name="rik_small")
