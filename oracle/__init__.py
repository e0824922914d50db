"""Python binding of the CPU oracle (oracle.c) and of the reference shim.

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg, never by the
``strassen_amd`` package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_lib = None
_ref = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_block_norm": (D, [P, I64]),
            "or_bsm_create": (P, [I32, P, P, I32, P, P, D]),
            "or_bsm_create_uniform": (P, [I64, I64, I32, I32, D]),
            "or_bsm_clone": (P, [P]),
            "or_bsm_view_reshape": (P, [P, I32, P, P, I32, P, P]),
            "or_bsm_destroy": (None, [P]),
            "or_bsm_nrow": (I64, [P]),
            "or_bsm_ncol": (I64, [P]),
            "or_bsm_nrb": (I32, [P]),
            "or_bsm_ncb": (I32, [P]),
            "or_bsm_from_dense": (C.c_int, [P, P, I64]),
            "or_bsm_to_dense": (C.c_int, [P, P, I64]),
            "or_bsm_get_meta": (C.c_int, [P, P, P]),
            "or_bsm_set_block": (C.c_int, [P, I32, I32, P]),
            "or_bsm_fill": (None, [P, D]),
            "or_bsm_zero": (None, [P]),
            "or_bsm_scale": (None, [P, D]),
            "or_bsm_calc_norms": (None, [P]),
            "or_bsm_recompress": (None, [P]),
            "or_bsm_add": (C.c_int, [P, P, D]),
            "or_bsm_add_transpose": (C.c_int, [P, P]),
            "or_bsm_frobenius_norm": (D, [P]),
            "or_bsm_dot": (D, [P, P, D]),
            "or_bsm_nalloc": (I64, [P]),
            "or_matmult": (C.c_int, [P, P, C.c_int, P, C.c_int, D, D, D, P, I64, P]),
            "or_strassen_dryrun": (C.c_int, [P, P, D, P, I64, P, P]),
            "or_matmult_strassen_sparse": (C.c_int, [P, P, P, D, I64, P, I64, P, P]),
            "or_matmult_strassen_tree": (C.c_int, [P, P, P, D, I64, P, P, I64, P, P]),
            "or_matmult_strassen_plan": (C.c_int, [P, P, P, D, I64, P, P, P, I64, P, P, P]),
            "or_dgemm": (None, [C.c_int, C.c_int, I64, I64, I64, D, P, I64, P, I64, D, P, I64]),
            "or_occ_ri_k": (C.c_int, [I64, I64, I64, I64, P, P, P, I32, D, P]),
            "or_occ_ri_k_dense": (C.c_int, [I64, I64, I64, I64, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own templates behind oracle/ref_shim.cpp (oracle/_ref)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref not built")
        R = C.CDLL(REF_PATH, mode=C.RTLD_GLOBAL)
        R.ref_norm.restype = D
        R.ref_norm.argtypes = [P, I64]
        R.ref_set_threads.restype = C.c_int
        R.ref_set_threads.argtypes = [C.c_int]
        R.ref_from_dense.restype = C.c_int
        R.ref_from_dense.argtypes = [P, I64, I64, I32, I32, D, P, P]
        R.ref_matmult.restype = C.c_int
        R.ref_matmult.argtypes = [P, I64, I64, D, P, I64, I64, D, P, I64, I64, D, C.c_int,
                                  I32, C.c_int, C.c_int, D, D, D, P, P, P]
        R.ref_dryrun.restype = C.c_int
        R.ref_dryrun.argtypes = [P, P, I64, I32, D, D, P, I64]
        R.ref_strassen.restype = C.c_int
        R.ref_strassen.argtypes = [P, P, P, I64, I32, D, D, I64, P]
        R.ref_occ_ri_k.restype = C.c_int
        R.ref_occ_ri_k.argtypes = [I64, I64, I64, I64, P, P, P, I32, D, P]
        R.ref_scale_recompress.restype = C.c_int
        R.ref_scale_recompress.argtypes = [P, I64, I64, I32, I32, D, D, P, P, P]
        R.ref_dot.restype = D
        R.ref_dot.argtypes = [P, P, I64, I64, I32, I32, D, D]
        _ref = R
    return _ref


def block_norm(x) -> float:
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F"))
    return lib().or_block_norm(_ptr(x), x.size)


class OBsm:
    """Oracle block-sparse matrix (host, C)."""

    def __init__(self, rsz, csz, thresh=0.0, rscr=None, cscr=None, _handle=None):
        L = lib()
        if _handle is not None:
            self.h = _handle
        else:
            rsz = np.ascontiguousarray(rsz, dtype=np.int32)
            csz = np.ascontiguousarray(csz, dtype=np.int32)
            rs = None if rscr is None else np.ascontiguousarray(rscr, dtype=np.int32)
            cs = None if cscr is None else np.ascontiguousarray(cscr, dtype=np.int32)
            self.h = L.or_bsm_create(len(rsz), _ptr(rsz), None if rs is None else _ptr(rs),
                                     len(csz), _ptr(csz), None if cs is None else _ptr(cs), thresh)
        self.nrow = L.or_bsm_nrow(self.h)
        self.ncol = L.or_bsm_ncol(self.h)
        self.nrb = L.or_bsm_nrb(self.h)
        self.ncb = L.or_bsm_ncb(self.h)

    @classmethod
    def uniform(cls, nrow, ncol, bs_r, bs_c, thresh=0.0):
        return cls(None, None, _handle=lib().or_bsm_create_uniform(nrow, ncol, bs_r, bs_c, thresh))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_bsm_destroy(self.h)
            self.h = None

    def clone(self):
        return OBsm(None, None, _handle=lib().or_bsm_clone(self.h))

    def view_reshape(self, rsz, csz, rscr=None, cscr=None):
        r = np.ascontiguousarray(rsz, dtype=np.int32)
        c = np.ascontiguousarray(csz, dtype=np.int32)
        rs = None if rscr is None else np.ascontiguousarray(rscr, dtype=np.int32)
        cs = None if cscr is None else np.ascontiguousarray(cscr, dtype=np.int32)
        h = lib().or_bsm_view_reshape(self.h, len(r), _ptr(r), None if rs is None else _ptr(rs), len(c), _ptr(c),
                                      None if cs is None else _ptr(cs))
        if not h:
            raise ValueError("view_reshape: geometry does not fit")
        return OBsm(None, None, _handle=h)

    def from_dense(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == (self.nrow, self.ncol)
        rc = lib().or_bsm_from_dense(self.h, _ptr(a), self.nrow)
        assert rc == 0
        return self

    def to_dense(self):
        out = np.zeros((self.nrow, self.ncol), dtype=np.float64, order="F")
        lib().or_bsm_to_dense(self.h, _ptr(out), self.nrow)
        return out

    def meta(self):
        nb = self.nrb * self.ncb
        pres = np.zeros(nb, dtype=np.uint8)
        norms = np.zeros(nb, dtype=np.float64)
        lib().or_bsm_get_meta(self.h, _ptr(pres), _ptr(norms))
        return (pres.reshape(self.nrb, self.ncb, order="F").astype(bool),
                norms.reshape(self.nrb, self.ncb, order="F"))

    def fill(self, v):
        lib().or_bsm_fill(self.h, v)

    def zero(self):
        lib().or_bsm_zero(self.h)

    def scale(self, s):
        lib().or_bsm_scale(self.h, s)

    def calc_norms(self):
        lib().or_bsm_calc_norms(self.h)

    def recompress(self):
        lib().or_bsm_recompress(self.h)

    def add(self, rhs, sign=1.0):
        assert lib().or_bsm_add(self.h, rhs.h, sign) == 0

    def add_transpose(self, rhs):
        assert lib().or_bsm_add_transpose(self.h, rhs.h) == 0

    def frobenius_norm(self):
        return lib().or_bsm_frobenius_norm(self.h)

    def dot(self, other, thresh=0.0):
        return lib().or_bsm_dot(self.h, other.h, thresh)

    def nalloc(self):
        return lib().or_bsm_nalloc(self.h)


def matmult(Cm, A, transA, B, transB, thresh, alpha=1.0, beta=0.0, cap=None):
    """Screened multiply; returns the task list (n, 3) int32 of (ib, jb, kb)."""
    if cap is None:
        cap = int(Cm.nrb) * int(Cm.ncb) * int(A.ncb if not transA else A.nrb)
    tasks = np.zeros((max(cap, 1), 3), dtype=np.int32)
    nt = I64(0)
    rc = lib().or_matmult(Cm.h, A.h, int(transA), B.h, int(transB), thresh, alpha, beta,
                          _ptr(tasks), cap, C.byref(nt))
    if rc != 0:
        raise ValueError(f"or_matmult failed: {rc}")
    return tasks[: min(nt.value, cap)].copy()


def strassen_dryrun(A, B, thresh_per_block):
    nib = A.nrb
    max_level = int(np.log2(nib + 0.5))
    sizes = []
    for lvl in range(max_level):
        sb = 1 << (lvl + 1)
        nsb = (nib + sb - 1) // sb
        sizes.append(nsb ** 3)
    total = max(sum(sizes), 1)
    bits = np.zeros(total, dtype=np.uint8)
    lsz = np.zeros(max_level + 1, dtype=np.int64)
    lfl = np.zeros(max_level + 1, dtype=np.uint64)
    rc = lib().or_strassen_dryrun(A.h, B.h, thresh_per_block, _ptr(bits), total, _ptr(lsz), _ptr(lfl))
    if rc < 0:
        raise ValueError(f"dryrun failed {rc}")
    out, off = [], 0
    for s in sizes:
        out.append(bits[off:off + s].copy())
        off += s
    return out, lfl


def tree_level_sizes(nib):
    max_level = int(np.log2(nib + 0.5))
    return [((nib + (1 << (l + 1)) - 1) >> (l + 1)) ** 3 for l in range(max_level)]


def forced_tree(nib, levels):
    """Decision tree with every node of the given levels set (edge nodes are
    the caller's responsibility)."""
    return [np.full(s, 1 if l in levels else 0, np.uint8) for l, s in enumerate(tree_level_sizes(nib))]


def _flat_tree(tree):
    return np.ascontiguousarray(np.concatenate(tree) if len(tree) else np.zeros(1), dtype=np.uint8)


def strassen_sparse(Cm, A, B, thresh, cutoff, tree=None, dense=None, counts=False):
    """Hybrid multiply with the dry-run tree, or caller Strassen bits `tree` and
    dense-root bits `dense` (per-level lists in dry-run layout).  Returns the
    leaf task list and the number of Strassen roots (and dense roots when
    counts=True)."""
    cap = int(A.nrb) ** 3
    leaf = np.zeros((max(cap, 1), 3), dtype=np.int32)
    nl = I64(0)
    nr = I64(0)
    nd = I64(0)
    tb = None if tree is None else _flat_tree(tree)
    db = None if dense is None else _flat_tree(dense)
    if tb is None and db is not None:
        tb = np.zeros_like(db)
    rc = lib().or_matmult_strassen_plan(Cm.h, A.h, B.h, thresh, cutoff, None if tb is None else _ptr(tb),
                                        None if db is None else _ptr(db), _ptr(leaf), cap, C.byref(nl),
                                        C.byref(nr), C.byref(nd))
    if rc != 0:
        raise ValueError(f"strassen failed {rc}")
    if counts:
        return leaf[: nl.value].copy(), nr.value, nd.value
    return leaf[: nl.value].copy(), nr.value


def occ_ri_k(n_bf, n_aux, n_occ, n_vec2, ints3c, mo, keys, bs, thresh):
    K = np.zeros((n_bf, n_occ), dtype=np.float64, order="F")
    i3 = np.asfortranarray(ints3c, dtype=np.float64)
    m = np.asfortranarray(mo, dtype=np.float64)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    rc = lib().or_occ_ri_k(n_bf, n_aux, n_occ, n_vec2, _ptr(i3), _ptr(m), _ptr(k), bs, thresh, _ptr(K))
    assert rc == 0
    return K


def occ_ri_k_dense(n_bf, n_aux, n_occ, n_vec2, ints3c, mo, keys):
    K = np.zeros((n_bf, n_occ), dtype=np.float64, order="F")
    i3 = np.asfortranarray(ints3c, dtype=np.float64)
    m = np.asfortranarray(mo, dtype=np.float64)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    rc = lib().or_occ_ri_k_dense(n_bf, n_aux, n_occ, n_vec2, _ptr(i3), _ptr(m), _ptr(k), _ptr(K))
    assert rc == 0
    return K
