"""Hand-designed block-sparse inputs whose expected results are fixed by
construction (not computed by any implementation under test).

Grid: 11 x 10 elements in target 4 x 4 blocks, so block rows hold 4, 4, 3
rows and block columns 4, 4, 2 columns; block b = rb + 3 cb.  Every block has
a single non-zero entry, so its Frobenius norm is |x| exactly, and all values
and thresholds are exact binary fractions.  The per-block thresholds of the
reference use the TARGET block size 4 x 4 = 16 (blocksparsematrix.hpp:150,
:273) even for the ragged blocks of the last block row / column, whose actual
sizes are 12, 8 and 6: several blocks sit exactly at the target threshold
(kept: the tests are `norm < thr` drops / `est >= thr` keeps) or exactly at
what an actual-size threshold would be, so a `<=`/`<` slip or an actual-size
threshold changes the result.
"""
import numpy as np

NR, NC, BS = 11, 10, 4
ROW0 = [0, 4, 8]
COL0 = [0, 4, 8]


def _single_entry(values):
    """Dense 11 x 10 array with values[b] at the first element of block b."""
    a = np.zeros((NR, NC), order="F")
    for b, v in enumerate(values):
        rb, cb = b % 3, b // 3
        a[ROW0[rb], COL0[cb]] = v
    return a


# --- scale(0.5) then recompress(), BSM threshold 2^-4 (per block: 1.0) ------
# Block values before scaling.  Copy-in drops |x| < 1 (block 4).  scale(0.5)
# multiplies each cached norm by 0.25 (the reference's s*s rule,
# blocksparsematrix.hpp:71-77), and recompress() then drops cached norms below
# 1.0: kept are the blocks with |x| >= 4.
RECOMPRESS_THRESH = 2.0 ** -4
RECOMPRESS_SCALE = 0.5
RECOMPRESS_VALUES = [4.0,   # b0 16 elements: cached 1.0, exactly at the threshold -> kept
                     3.5,   # b1 16: 0.875 -> dropped
                     3.0,   # b2 12 (ragged row): 0.75 -> dropped (an actual-size threshold 0.75 would keep it)
                     8.0,   # b3 16: 2.0 -> kept
                     0.5,   # b4 16: dropped at copy-in (0.5 < 1.0)
                     4.5,   # b5 12: 1.125 -> kept
                     2.0,   # b6 8 (ragged column): 0.5 -> dropped (actual-size threshold 0.5 would keep it)
                     -6.0,  # b7 8: 1.5 -> kept
                     2.0]   # b8 6: 0.5 -> dropped (actual-size threshold 0.375 would keep it)
RECOMPRESS_KEPT = [0, 3, 5, 7]


def recompress_case():
    """(dense input, threshold, scale, expected presence (3 x 3, grid order),
    expected cached norms of the kept blocks)."""
    a = _single_entry(RECOMPRESS_VALUES)
    pres = np.zeros(9, dtype=bool)
    pres[RECOMPRESS_KEPT] = True
    norms = np.zeros(9)
    for b in RECOMPRESS_KEPT:
        norms[b] = abs(RECOMPRESS_VALUES[b]) * RECOMPRESS_SCALE ** 2
    return (a, RECOMPRESS_THRESH, RECOMPRESS_SCALE, pres.reshape(3, 3, order="F"),
            norms.reshape(3, 3, order="F"))


# --- dot(A, B, 2^-4) (per block: 1.0), both built at threshold 0 -------------
# A holds 1.0 in every block, so block b's norm product and its contribution
# are both c_b = B's value.
DOT_THRESH = 2.0 ** -4
DOT_CONTRIB = [1.0,     # b0 16: exactly at the threshold -> included
               -0.875,  # b1 16: |c| < 1 -> excluded
               0.75,    # b2 12: excluded (actual-size threshold 0.75 would include it)
               -2.0,    # b3 -> included
               4.0,     # b4 -> included
               8.0,     # b5 12 -> included
               0.5,     # b6 8: excluded (actual-size 0.5 would include it)
               -16.0,   # b7 8 -> included
               0.4375]  # b8 6: excluded (actual-size 0.375 would include it)
DOT_EXPECTED = 1.0 - 2.0 + 4.0 + 8.0 - 16.0      # -5.0
DOT_EXPECTED_UNSCREENED = float(sum(DOT_CONTRIB))  # threshold 0: every block


def dot_case():
    a = _single_entry([1.0] * 9)
    b = _single_entry(DOT_CONTRIB)
    return a, b, DOT_THRESH, DOT_EXPECTED, DOT_EXPECTED_UNSCREENED


def near_tie_case(nb=6, bs=16, seed=3):
    """Screening at near-ties: every block of A and B is a permutation of one
    fixed bs*bs-value vector, so all block norms are equal in exact arithmetic
    and differ only by the rounding of their summation order.  Returns dense
    A, B and bs; callers put the threshold at the median of the candidate
    triples' estimates, so every triple sits within a few units in the last
    place of est/thr = 1 and both outcomes occur."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(bs * bs) * np.exp(rng.uniform(-2.0, 2.0, bs * bs))
    n = nb * bs

    def build():
        a = np.zeros((n, n), order="F")
        for cb in range(nb):
            for rb in range(nb):
                a[rb * bs:(rb + 1) * bs, cb * bs:(cb + 1) * bs] = rng.permutation(v).reshape(bs, bs, order="F")
        return a

    return build(), build(), bs
