"""strassen_amd -- MI355X block-sparse multiply (diersam/Strassen hot path).

This module is the Python view of the C ABI declared in include/strassen_amd.h
(used by the tests and bench.py).  It loads the in-tree HIP library
strassen_amd/lib/libstrassen_amd.so and fails loudly when it is missing: there
is no host fallback.  The C++ drop-in layer for callers of the reference's
headers lives in strassen_amd/cpp/.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libstrassen_amd.so")

SA_OK = 0
MODE_FP64, MODE_OZAKI = 0, 1
OK, ERR_INVALID, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED = 0, -1, -2, -3, -4  # SA_OK, SA_ERR_*
SLICE_FP16, SLICE_BF16, SLICE_INT8 = 0, 1, 2
SCHED_HEURISTIC, SCHED_MLP, SCHED_TREE = 0, 1, 2

_lib = None
_live_ctx = weakref.WeakSet()
_live_bsm = weakref.WeakSet()


@atexit.register
def _teardown():
    """Release device objects before interpreter shutdown: matrices first,
    then their contexts (a matrix refers to its context's stream)."""
    for m in list(_live_bsm):
        m.close()
    for c in list(_live_ctx):
        c.close()


P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double
INT = C.c_int


class Opts(C.Structure):
    _fields_ = [("mode", I32), ("slices", I32), ("slice_type", I32), ("tile", I32),
                ("scheduler", I32), ("shared_exp", I32), ("cutoff", I64), ("tree", P), ("tree_dense", P)]


# name -> (restype, argtypes); every function declared in include/strassen_amd.h
ABI = {
    "sa_opts_default": (None, [P]),
    "sa_last_error": (C.c_char_p, []),
    "sa_device_count": (INT, [P]),
    "sa_ctx_create": (INT, [INT, P]),
    "sa_ctx_destroy": (INT, [P]),
    "sa_ctx_synchronize": (INT, [P]),
    "sa_ctx_set_profiling": (INT, [P, INT]),
    "sa_ctx_profile": (INT, [P, C.c_char_p, P, P]),
    "sa_ctx_profile_reset": (INT, [P]),
    "sa_ctx_gemm_config": (INT, [P, P, C.c_char_p, I32]),
    "sa_ctx_kwindows": (INT, [P, P]),
    "sa_ctx_graph_stats": (INT, [P, P, P, P]),
    "sa_ctx_slice_pairs": (INT, [P, P, P]),
    "sa_mlp_info": (INT, [P, P, P]),
    "sa_mlp_set": (INT, [P, P, I64, P, P]),
    "sa_mlp_get": (INT, [P, P, P, P]),
    "sa_mlp_load": (INT, [P, C.c_char_p]),
    "sa_mlp_forward": (INT, [P, P, I64, P]),
    "sa_mlp_grad": (INT, [P, P, P, P, I64, P, P]),
    "sa_mlp_train": (INT, [P, P, P, P, I64, I32, D, P]),
    "sa_strassen_features": (INT, [P, P, P, D, P, I64, P]),
    "sa_strassen_plan": (INT, [P, P, P, D, P, P, I64, P, P, P]),
    "sa_strassen_last_plan": (INT, [P, P, P, P]),
    "sa_ctx_mark": (INT, [P, INT]),
    "sa_ctx_elapsed": (INT, [P, INT, INT, P]),
    "sa_bsm_create": (INT, [P, I64, I64, I32, I32, D, P]),
    "sa_bsm_create_blocked": (INT, [P, I32, P, P, I32, P, P, D, P]),
    "sa_bsm_destroy": (INT, [P]),
    "sa_bsm_info": (INT, [P, P, P, P, P, P, P]),
    "sa_bsm_copy": (INT, [P, P]),
    "sa_bsm_from_dense": (INT, [P, P, I64]),
    "sa_bsm_to_dense": (INT, [P, P, I64]),
    "sa_bsm_get_meta": (INT, [P, P, P]),
    "sa_bsm_generate": (INT, [P, C.c_uint64, P]),
    "sa_bsm_view_reshape": (INT, [P, I32, P, P, I32, P, P, P]),
    "sa_bsm_fill": (INT, [P, D]),
    "sa_bsm_zero": (INT, [P]),
    "sa_bsm_scale": (INT, [P, D]),
    "sa_bsm_scale_ex": (INT, [P, D, INT]),
    "sa_bsm_reserve": (INT, [P, I64]),
    "sa_bsm_get_blocks": (INT, [P, I32, I32, I32, I32, P, I64]),
    "sa_matmult_noalloc": (INT, [P, P, INT, P, INT, D, D, D]),
    "sa_matmult_scratch_bytes": (INT, [P]),
    "sa_bsm_calc_norms": (INT, [P]),
    "sa_bsm_recompress": (INT, [P]),
    "sa_bsm_add": (INT, [P, P, D]),
    "sa_bsm_add_transpose": (INT, [P, P]),
    "sa_bsm_frobenius_norm": (INT, [P, P]),
    "sa_bsm_dot": (INT, [P, P, D, P]),
    "sa_matmult": (INT, [P, P, INT, P, INT, D, D, D, P]),
    "sa_tasklist_get": (INT, [P, P, I64, P]),
    "sa_dgemm": (INT, [P, INT, INT, I64, I64, I64, D, P, I64, P, I64, D, P, I64]),
    "sa_dnrm2": (INT, [P, I64, P, P]),
    "sa_ddot": (INT, [P, I64, P, P, P]),
    "sa_daxpy": (INT, [P, I64, D, P, P]),
    "sa_inv_sqrt": (INT, [P, I64, P, P, P, I32, D, P, P]),
    "sa_ri_k_step": (INT, [P, P, P, P, I32, P, I32, P, D, P, P]),
    "sa_occ_ri_k": (INT, [P, I64, I64, I64, I64, P, P, P, I32, D, I32, P, P]),
    "sa_strassen_dryrun": (INT, [P, P, D, P, I64, P, P, P]),
    "sa_matmult_strassen_sparse": (INT, [P, P, INT, P, INT, D, D, D, P]),
    "sa_strassen_last_stats": (INT, [P, P, P, P, P, P, P]),
}


class SAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"strassen_amd error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != SA_OK:
        raise SAError(rc, lib().sa_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().sa_device_count(C.byref(n)))
    return n.value


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().sa_opts_default(C.byref(o))
    for k, v in kw.items():
        if k in ("tree", "tree_dense"):
            arr = np.ascontiguousarray(np.concatenate(v) if isinstance(v, (list, tuple)) else v, dtype=np.uint8)
            setattr(o, "_" + k + "_keep", arr)
            setattr(o, k, arr.ctypes.data)
        else:
            setattr(o, k, v)
    return o


class Context:
    def __init__(self, device: int = 0):
        self.h = P()
        _check(lib().sa_ctx_create(device, C.byref(self.h)))
        _live_ctx.add(self)

    def close(self):
        if self.h:
            lib().sa_ctx_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().sa_ctx_synchronize(self.h))

    def set_profiling(self, on: bool):
        _check(lib().sa_ctx_set_profiling(self.h, int(on)))

    def profile(self, cls: str):
        ms = D(0.0)
        n = I64(0)
        _check(lib().sa_ctx_profile(self.h, cls.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def profile_reset(self):
        _check(lib().sa_ctx_profile_reset(self.h))

    def gemm_config(self):
        cfg = I32(-1)
        buf = C.create_string_buffer(64)
        _check(lib().sa_ctx_gemm_config(self.h, C.byref(cfg), buf, 64))
        return cfg.value, buf.value.decode()

    def kwindows(self) -> int:
        """k-windows of the last screened multiply (1 = one grouped launch)."""
        n = I32(0)
        _check(lib().sa_ctx_kwindows(self.h, C.byref(n)))
        return n.value

    def graph_stats(self) -> dict:
        """HIP-graph counters: executables instantiated, updated in place, launches."""
        v = [C.c_int64(0) for _ in range(3)]
        _check(lib().sa_ctx_graph_stats(self.h, *(C.byref(x) for x in v)))
        return dict(zip(("instantiated", "updated", "launched"), (x.value for x in v)))

    def slice_pairs(self) -> tuple:
        """Split-precision slice-pair work since the last call (then reset):
        (multiplied, skipped as zero), in slice-pair k-steps x GEMM tiles."""
        v = [C.c_int64(0) for _ in range(2)]
        _check(lib().sa_ctx_slice_pairs(self.h, *(C.byref(x) for x in v)))
        return v[0].value, v[1].value

    def mark(self, slot: int):
        _check(lib().sa_ctx_mark(self.h, slot))

    def elapsed(self, s0: int, s1: int) -> float:
        ms = C.c_float(0.0)
        _check(lib().sa_ctx_elapsed(self.h, s0, s1, C.byref(ms)))
        return ms.value

    def tasklist(self, cap=None):
        n = I64(0)
        _check(lib().sa_tasklist_get(self.h, None, 0, C.byref(n)))
        cap = n.value if cap is None else cap
        out = np.zeros((max(cap, 1), 3), dtype=np.int32)
        _check(lib().sa_tasklist_get(self.h, _ptr(out), cap, C.byref(n)))
        return out[: min(cap, n.value)]


class BSM:
    """Device block-sparse matrix (include/strassen_amd.h sa_bsm_*)."""

    def __init__(self, ctx: Context, rsz, csz, thresh=0.0, rscr=None, cscr=None, _uniform=None):
        self.ctx = ctx
        self.h = P()
        if _uniform is not None:
            nrow, ncol, bsr, bsc = _uniform
            _check(lib().sa_bsm_create(ctx.h, nrow, ncol, bsr, bsc, thresh, C.byref(self.h)))
        else:
            r = np.ascontiguousarray(rsz, dtype=np.int32)
            c = np.ascontiguousarray(csz, dtype=np.int32)
            rs = None if rscr is None else np.ascontiguousarray(rscr, dtype=np.int32)
            cs = None if cscr is None else np.ascontiguousarray(cscr, dtype=np.int32)
            _check(lib().sa_bsm_create_blocked(ctx.h, len(r), _ptr(r), None if rs is None else _ptr(rs),
                                               len(c), _ptr(c), None if cs is None else _ptr(cs), thresh,
                                               C.byref(self.h)))
        _live_bsm.add(self)
        self.nrow, self.ncol, self.nrb, self.ncb, _, _ = self.info(False)
        if _uniform is not None:  # target block size, ragged last block
            self.rsz = np.array([min(bsr, nrow - i * bsr) for i in range(self.nrb)], dtype=np.int64)
            self.csz = np.array([min(bsc, ncol - j * bsc) for j in range(self.ncb)], dtype=np.int64)
        else:
            self.rsz, self.csz = np.asarray(rsz, dtype=np.int64), np.asarray(csz, dtype=np.int64)

    @classmethod
    def uniform(cls, ctx, nrow, ncol, bs_r, bs_c, thresh=0.0):
        return cls(ctx, None, None, thresh, _uniform=(nrow, ncol, bs_r, bs_c))

    def close(self):
        if self.h:
            lib().sa_bsm_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, with_nalloc=True):
        nr, nc, nalloc, ab = I64(0), I64(0), I64(0), I64(0)
        rb, cb = I32(0), I32(0)
        _check(lib().sa_bsm_info(self.h, C.byref(nr), C.byref(nc), C.byref(rb), C.byref(cb),
                                 C.byref(nalloc) if with_nalloc else None, C.byref(ab)))
        return nr.value, nc.value, rb.value, cb.value, nalloc.value, ab.value

    def nalloc(self):
        return self.info(True)[4]

    def copy_from(self, src):
        _check(lib().sa_bsm_copy(self.h, src.h))
        return self

    def from_dense(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == (self.nrow, self.ncol), (a.shape, self.nrow, self.ncol)
        _check(lib().sa_bsm_from_dense(self.h, _ptr(a), self.nrow))
        return self

    def to_dense(self):
        out = np.zeros((self.nrow, self.ncol), dtype=np.float64, order="F")
        _check(lib().sa_bsm_to_dense(self.h, _ptr(out), self.nrow))
        return out

    def blocks(self, rb0, rb1, cb0, cb1):
        """Dense values of block rows [rb0, rb1) x block columns [cb0, cb1)
        (absent blocks zero); the reference's block(rb, cb) for one block."""
        rows = int(np.sum(self.rsz[rb0:rb1]))
        cols = int(np.sum(self.csz[cb0:cb1]))
        out = np.zeros((rows, cols), order="F")
        if rows and cols:
            _check(lib().sa_bsm_get_blocks(self.h, rb0, rb1, cb0, cb1, _ptr(out), rows))
        return out

    def meta(self):
        nb = self.nrb * self.ncb
        pres = np.zeros(max(nb, 1), dtype=np.uint8)
        norms = np.zeros(max(nb, 1), dtype=np.float64)
        _check(lib().sa_bsm_get_meta(self.h, _ptr(pres), _ptr(norms)))
        return (pres[:nb].reshape(self.nrb, self.ncb, order="F").astype(bool),
                norms[:nb].reshape(self.nrb, self.ncb, order="F"))

    def generate(self, seed, block_scale):
        s = np.asfortranarray(block_scale, dtype=np.float64)
        assert s.shape == (self.nrb, self.ncb)
        flat = np.ascontiguousarray(s.ravel(order="F"))
        _check(lib().sa_bsm_generate(self.h, seed, _ptr(flat)))
        return self

    def view_reshape(self, rsz, csz, rscr=None, cscr=None):
        """Read-only view with new block geometry over the same block payloads."""
        r = np.ascontiguousarray(rsz, dtype=np.int32)
        c = np.ascontiguousarray(csz, dtype=np.int32)
        rs = None if rscr is None else np.ascontiguousarray(rscr, dtype=np.int32)
        cs = None if cscr is None else np.ascontiguousarray(cscr, dtype=np.int32)
        v = BSM.__new__(BSM)
        v.ctx = self.ctx
        v.h = P()
        v._base = self  # keep the base alive
        _check(lib().sa_bsm_view_reshape(self.h, len(r), _ptr(r), None if rs is None else _ptr(rs), len(c), _ptr(c),
                                         None if cs is None else _ptr(cs), C.byref(v.h)))
        _live_bsm.add(v)
        v.nrow, v.ncol, v.nrb, v.ncb, _, _ = v.info(False)
        return v

    def fill(self, v):
        _check(lib().sa_bsm_fill(self.h, v))

    def zero(self):
        _check(lib().sa_bsm_zero(self.h))

    def scale(self, s, recompute_norms=False):
        """scale (blocksparsematrix.hpp:71-77: cached norms *= s*s), or with
        recompute_norms the cached norms recomputed from the scaled blocks."""
        if recompute_norms:
            _check(lib().sa_bsm_scale_ex(self.h, s, 1))
        else:
            _check(lib().sa_bsm_scale(self.h, s))

    def calc_norms(self):
        _check(lib().sa_bsm_calc_norms(self.h))

    def recompress(self):
        _check(lib().sa_bsm_recompress(self.h))

    def add(self, rhs, sign=1.0):
        _check(lib().sa_bsm_add(self.h, rhs.h, sign))

    def add_transpose(self, rhs):
        _check(lib().sa_bsm_add_transpose(self.h, rhs.h))

    def frobenius_norm(self):
        out = D(0.0)
        _check(lib().sa_bsm_frobenius_norm(self.h, C.byref(out)))
        return out.value

    def dot(self, other, thresh=0.0):
        out = D(0.0)
        _check(lib().sa_bsm_dot(self.h, other.h, thresh, C.byref(out)))
        return out.value


def matmult(Cm: BSM, A: BSM, transA, B: BSM, transB, thresh, alpha=1.0, beta=0.0, opts: Opts | None = None):
    _check(lib().sa_matmult(Cm.h, A.h, int(transA), B.h, int(transB), thresh, alpha, beta,
                            None if opts is None else C.byref(opts)))


def tree_level_sizes(nb):
    max_level = int(np.log2(nb + 0.5))
    return [((nb + (1 << (l + 1)) - 1) >> (l + 1)) ** 3 for l in range(max_level)]


def strassen_dryrun(A: BSM, B: BSM, thresh_per_block):
    """Device dry run: list of per-level bit arrays and the per-level flops."""
    sizes = tree_level_sizes(A.nrb)
    total = max(sum(sizes), 1)
    bits = np.zeros(total, dtype=np.uint8)
    lsz = np.zeros(len(sizes) + 1, dtype=np.int64)
    lfl = np.zeros(len(sizes) + 1, dtype=np.uint64)
    nl = I32(0)
    _check(lib().sa_strassen_dryrun(A.h, B.h, thresh_per_block, _ptr(bits), total, _ptr(lsz), _ptr(lfl),
                                    C.byref(nl)))
    out, off = [], 0
    for s in sizes:
        out.append(bits[off:off + s].copy())
        off += s
    return out, lfl


def matmult_strassen_sparse(Cm: BSM, A: BSM, B: BSM, thresh, cutoff=0, opts: Opts | None = None, tree=None,
                            dense=None):
    """Hybrid Strassen-sparse multiply (strassen.hpp:687-753).  tree / dense:
    caller Strassen-root / dense-root bits (per-level lists, dry-run layout)."""
    o = opts if opts is not None else default_opts()
    o.cutoff = int(cutoff)
    for name, bits in (("tree", tree), ("tree_dense", dense)):
        if bits is not None:
            arr = np.ascontiguousarray(np.concatenate(bits) if len(bits) else np.zeros(1), dtype=np.uint8)
            setattr(o, "_" + name + "_keep", arr)
            setattr(o, name, arr.ctypes.data)
            o.scheduler = SCHED_TREE
    _check(lib().sa_matmult_strassen_sparse(Cm.h, A.h, 0, B.h, 0, thresh, 1.0, 0.0, C.byref(o)))


# ------------------------------------------------------ learned path selector --
DEFAULT_PATH_MODEL = os.path.join(os.path.dirname(__file__), "data", "path_mlp.txt")


def mlp_info():
    """(features, outputs, parameters) of the path-selector network."""
    f, o, p = I32(0), I32(0), I64(0)
    _check(lib().sa_mlp_info(C.byref(f), C.byref(o), C.byref(p)))
    return f.value, o.value, p.value


def _rows(x, ncol):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, ncol))


def mlp_set(ctx: Context, params, mu=None, sd=None):
    f, _, npar = mlp_info()
    prm = np.ascontiguousarray(params, dtype=np.float64)
    assert prm.size == npar
    mu = np.ascontiguousarray(np.zeros(f) if mu is None else mu, dtype=np.float64)
    sd = np.ascontiguousarray(np.ones(f) if sd is None else sd, dtype=np.float64)
    _check(lib().sa_mlp_set(ctx.h, _ptr(prm), npar, _ptr(mu), _ptr(sd)))


def mlp_get(ctx: Context):
    f, _, npar = mlp_info()
    prm, mu, sd = np.zeros(npar), np.zeros(f), np.zeros(f)
    _check(lib().sa_mlp_get(ctx.h, _ptr(prm), _ptr(mu), _ptr(sd)))
    return prm, mu, sd


def mlp_load(ctx: Context, path: str = DEFAULT_PATH_MODEL):
    _check(lib().sa_mlp_load(ctx.h, path.encode()))


def mlp_save(ctx: Context, path: str):
    f, o, npar = mlp_info()
    prm, mu, sd = mlp_get(ctx)
    with open(path, "w") as fh:
        fh.write(f"path_mlp {f} 32 {o} {npar}\n")
        for v in np.concatenate([prm, mu, sd]):
            fh.write(repr(float(v)) + "\n")


def mlp_forward(ctx: Context, X):
    f, o, _ = mlp_info()
    X = _rows(X, f)
    Y = np.zeros((X.shape[0], o))
    _check(lib().sa_mlp_forward(ctx.h, _ptr(X), X.shape[0], _ptr(Y)))
    return Y


def mlp_grad(ctx: Context, X, T, M=None):
    f, o, npar = mlp_info()
    X, T = _rows(X, f), _rows(T, o)
    M = None if M is None else _rows(M, o)
    g = np.zeros(npar)
    loss = D(0.0)
    _check(lib().sa_mlp_grad(ctx.h, _ptr(X), _ptr(T), None if M is None else _ptr(M), X.shape[0], _ptr(g),
                             C.byref(loss)))
    return g, loss.value


def mlp_train(ctx: Context, X, T, M=None, epochs=1000, lr=3e-3):
    f, o, _ = mlp_info()
    X, T = _rows(X, f), _rows(T, o)
    M = None if M is None else _rows(M, o)
    loss = D(0.0)
    _check(lib().sa_mlp_train(ctx.h, _ptr(X), _ptr(T), None if M is None else _ptr(M), X.shape[0], int(epochs),
                              float(lr), C.byref(loss)))
    return loss.value


def strassen_features(Cm: BSM | None, A: BSM, B: BSM, thresh):
    """Node features of every dry-run node (rows in dry-run order)."""
    f, _, _ = mlp_info()
    nn = I64(0)
    _check(lib().sa_strassen_features(None if Cm is None else Cm.h, A.h, B.h, thresh, None, 0, C.byref(nn)))
    X = np.zeros((max(nn.value, 1), f))
    _check(lib().sa_strassen_features(None if Cm is None else Cm.h, A.h, B.h, thresh, _ptr(X), X.size,
                                      C.byref(nn)))
    return X[: nn.value]


def _split_levels(flat, nb):
    out, off = [], 0
    for s in tree_level_sizes(nb):
        out.append(flat[off:off + s].copy())
        off += s
    return out


def strassen_plan(Cm: BSM | None, A: BSM, B: BSM, thresh):
    """Learned plan: (Strassen bits, dense bits) per level, overruled flag,
    model flops of the plan and of the heuristic plan."""
    total = max(sum(tree_level_sizes(A.nrb)), 1)
    sb = np.zeros(total, dtype=np.uint8)
    db = np.zeros(total, dtype=np.uint8)
    ov = I32(0)
    mc, hc = D(0.0), D(0.0)
    _check(lib().sa_strassen_plan(None if Cm is None else Cm.h, A.h, B.h, thresh, _ptr(sb), _ptr(db), total,
                                  C.byref(ov), C.byref(mc), C.byref(hc)))
    return _split_levels(sb, A.nrb), _split_levels(db, A.nrb), bool(ov.value), mc.value, hc.value


def strassen_last_plan():
    s, ov = I32(0), I32(0)
    mc, hc = D(0.0), D(0.0)
    _check(lib().sa_strassen_last_plan(C.byref(s), C.byref(ov), C.byref(mc), C.byref(hc)))
    return {"scheduler": s.value, "overruled": bool(ov.value), "model_cost": mc.value, "heuristic_cost": hc.value}


def strassen_last_stats():
    v = [I64(0) for _ in range(7)]
    _check(lib().sa_strassen_last_stats(*[C.byref(x) for x in v]))
    keys = ["nroots", "nnodes", "nleaf_products", "scratch_bytes", "batches", "screened", "dense_roots"]
    return {k: x.value for k, x in zip(keys, v)}


def inv_sqrt(ctx: Context, M, X_bsm: BSM | None = None, max_iter=100, tol=1e-14, want_x=True):
    """X = M^{-1/2} on the device (coupled Newton-Schulz).  Returns (X or None,
    iterations, residual ||XMX - I||_F/sqrt(n))."""
    M = np.asfortranarray(M, dtype=np.float64)
    n = M.shape[0]
    X = np.zeros((n, n), dtype=np.float64, order="F") if want_x else None
    it = I32(0)
    res = D(0.0)
    _check(lib().sa_inv_sqrt(ctx.h, n, _ptr(M), None if X is None else _ptr(X), None if X_bsm is None else X_bsm.h,
                             max_iter, tol, C.byref(it), C.byref(res)))
    return X, it.value, res.value


def occ_ri_k(ctx: Context, n_bf, n_aux, n_occ, n_vec2, ints3c, mo, keys, bs, thresh, teams=0,
             opts: Opts | None = None):
    """Device occupied RI-K (sa_occ_ri_k): compressed integrals (n_vec2 x n_aux),
    MO coefficients (n_bf x n_occ) and v2m keys in, K (n_bf x n_occ) out;
    teams = members of the P loop's team (0: 8)."""
    i3 = np.asfortranarray(ints3c, dtype=np.float64)
    m = np.asfortranarray(mo, dtype=np.float64)
    k = np.ascontiguousarray(np.ravel(keys), dtype=np.uint64)
    assert i3.shape == (n_vec2, n_aux) and m.shape == (n_bf, n_occ) and k.size == n_vec2
    K = np.zeros((n_bf, n_occ), order="F")
    _check(lib().sa_occ_ri_k(ctx.h, n_bf, n_aux, n_occ, n_vec2, _ptr(i3), _ptr(m), _ptr(k), bs, thresh, teams,
                             None if opts is None else C.byref(opts), _ptr(K)))
    return K


def ri_k_step(X: BSM, B: BSM, Bp: BSM, K: BSM, mu_sizes, i_chunks, thresh, opts_t: Opts | None = None,
              opts_k: Opts | None = None):
    mu = np.ascontiguousarray(mu_sizes, dtype=np.int32)
    ic = np.ascontiguousarray(i_chunks, dtype=np.int32)
    _check(lib().sa_ri_k_step(X.h, B.h, Bp.h, K.h, len(mu), _ptr(mu), len(ic), _ptr(ic), thresh,
                              None if opts_t is None else C.byref(opts_t), None if opts_k is None else C.byref(opts_k)))
