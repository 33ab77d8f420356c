"""ctypes access to the CPU oracle (oracle/_build/liboracle_ed.so).

Test infrastructure: imported by tests/, the smoke check and the CPU-baseline
leg of bench.py only.  The oracle is built by ``make -C oracle`` (called from
``__graft_entry__.build()``).
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle_ed.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, u64, f64 = ctypes.c_int, ctypes.c_long, ctypes.c_uint64, ctypes.c_double
        L.oracle_lattice.argtypes = [ctypes.c_char_p, i32, i32, i32, P, P, i32]
        L.oracle_lattice.restype = i32
        L.oracle_basis.argtypes = [i32, i32, P, i64]
        L.oracle_basis.restype = i64
        model = [i32] + [P] * 7  # n_bonds, bond_s, bond_t, jz, jxy, k, h, d
        L.oracle_build_csr.argtypes = [i32, i32] + model + [i64, i64, P, P, P, i64]
        L.oracle_build_csr.restype = i64
        L.oracle_spmv.argtypes = [i32, i32] + model + [P, P]
        L.oracle_spmv.restype = i32
        L.oracle_tridiag.argtypes = [i32, P, P, P, P]
        L.oracle_tridiag.restype = i32
        L.oracle_tmatrix.argtypes = [i32, P, P, P, P, P, P, P, P]
        L.oracle_tmatrix.restype = i32
        L.oracle_start_value.argtypes = [u64, u64, f64, f64]
        L.oracle_start_value.restype = f64
        L.oracle_sparsediag.argtypes = [i32, i32] + model + [i32, i32, u64, i32, P, P, P, P, i32, P, P, P, P]
        L.oracle_sparsediag.restype = i32
        L.oracle_bench_hv.argtypes = [i32, i32] + model + [i64, i32, i32, P, P]
        L.oracle_bench_hv.restype = f64
        L.oracle_sparsediag_timed.argtypes = [i32, i32] + model + [i32, i32, u64, i32, P, P, P, P, P]
        L.oracle_sparsediag_timed.restype = i32
        L.oracle_ground_and_gap.argtypes = [i32, P, P, P, P]
        L.oracle_ground_and_gap.restype = i32
        L.oracle_expect_site_term.argtypes = [i32, i32, P, i32, i32, i32, i32, P]
        L.oracle_expect_site_term.restype = i32
        L.oracle_translation_momenta.argtypes = [ctypes.c_char_p, i32, i32, i32, i32, P, P]
        L.oracle_translation_momenta.restype = i32
        L.oracle_translations.argtypes = [ctypes.c_char_p, i32, i32, i32, i32, P, i32, P, P, P]
        L.oracle_translations.restype = i32
        L.oracle_bloch_basis.argtypes = [i32, i32, i32, P, P, P, i64, P, P]
        L.oracle_bloch_basis.restype = i64
        L.oracle_bloch_dense.argtypes = [i32, i32] + model + [i32, P, P, P, i64, P, P]
        L.oracle_bloch_dense.restype = i64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


@dataclass
class Model:
    """Spin-1/2 model on N sites: bonds (s, t) in edge order with per-bond
    Jz, Jxy, K and per-site h, D."""
    n_sites: int
    bond_s: np.ndarray
    bond_t: np.ndarray
    jz: np.ndarray
    jxy: np.ndarray
    k: np.ndarray = None
    h: np.ndarray = None
    d: np.ndarray = None
    bond_type: np.ndarray = field(default=None)

    def __post_init__(self):
        nb = len(self.bond_s)
        self.bond_s = np.ascontiguousarray(self.bond_s, dtype=np.int32)
        self.bond_t = np.ascontiguousarray(self.bond_t, dtype=np.int32)
        self.jz = np.ascontiguousarray(np.broadcast_to(self.jz, (nb,)), dtype=np.float64)
        self.jxy = np.ascontiguousarray(np.broadcast_to(self.jxy, (nb,)), dtype=np.float64)
        self.k = np.ascontiguousarray(np.broadcast_to(0.0 if self.k is None else self.k, (nb,)), dtype=np.float64)
        self.h = np.ascontiguousarray(np.broadcast_to(0.0 if self.h is None else self.h, (self.n_sites,)), dtype=np.float64)
        self.d = np.ascontiguousarray(np.broadcast_to(0.0 if self.d is None else self.d, (self.n_sites,)), dtype=np.float64)

    @property
    def n_bonds(self):
        return len(self.bond_s)

    def cargs(self):
        return [self.n_bonds, _p(self.bond_s), _p(self.bond_t), _p(self.jz), _p(self.jxy),
                _p(self.k), _p(self.h), _p(self.d)]


def lattice(name, L, W=0, H=0):
    n = ctypes.c_int(0)
    buf = np.zeros(3 * 4096, dtype=np.int32)
    nb = lib().oracle_lattice(name.encode(), L, W, H, ctypes.byref(n), _p(buf), 4096)
    if nb < 0:
        raise ValueError(f"oracle_lattice({name!r}) failed: {nb}")
    b = buf[: 3 * nb].reshape(nb, 3)
    return n.value, b[:, 0].copy(), b[:, 1].copy(), b[:, 2].copy()


def lattice_model(name, L, W=0, jz=1.0, jxy=1.0, h=0.0, by_type=None):
    """Model on a named lattice.  by_type maps bond type -> (Jz, Jxy)."""
    N, s, t, ty = lattice(name, L, W)
    jzv = np.full(len(s), float(jz))
    jxyv = np.full(len(s), float(jxy))
    if by_type:
        for typ, (a, b) in by_type.items():
            jzv[ty == typ] = a
            jxyv[ty == typ] = b
    return Model(N, s, t, jzv, jxyv, h=h, bond_type=ty)


def basis(N, n_up):
    from math import comb
    D = comb(N, n_up) if 0 <= n_up <= N else 0
    out = np.zeros(max(D, 1), dtype=np.uint32)
    got = lib().oracle_basis(N, n_up, _p(out), D)
    assert got == D
    return out[:D]


def build_csr(m: Model, n_up, row0=0, row1=0):
    from math import comb
    D = comb(m.n_sites, n_up)
    r1 = D if row1 <= 0 else row1
    nrows = r1 - row0
    cap = max(16, nrows * (m.n_bonds + 2))
    rowptr = np.zeros(nrows + 1, dtype=np.uint64)
    col = np.zeros(cap, dtype=np.uint32)
    val = np.zeros(cap, dtype=np.float64)
    nnz = lib().oracle_build_csr(m.n_sites, n_up, *m.cargs(), row0, r1, _p(rowptr), _p(col), _p(val), cap)
    if nnz < 0:
        raise RuntimeError(f"oracle_build_csr failed ({nnz})")
    return rowptr, col[:nnz].copy(), val[:nnz].copy()


def dense(m: Model, n_up):
    rowptr, col, val = build_csr(m, n_up)
    D = len(rowptr) - 1
    A = np.zeros((D, D))
    for i in range(D):
        for p in range(int(rowptr[i]), int(rowptr[i + 1])):
            A[i, col[p]] += val[p]
    return A


def spmv(m: Model, n_up, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    rc = lib().oracle_spmv(m.n_sites, n_up, *m.cargs(), _p(x), _p(y))
    assert rc == 0
    return y


def tridiag(d, e, vectors=False):
    d = np.ascontiguousarray(d, dtype=np.float64)
    n = len(d)
    e = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.float64)[: max(n - 1, 0)], [0.0]]))
    ev = np.zeros(n)
    z = np.zeros((n, n), order="F") if vectors else None
    rc = lib().oracle_tridiag(n, _p(d), _p(e), _p(ev), _p(z) if vectors else None)
    if rc != 0:
        raise RuntimeError(f"tridiag failed: {rc}")
    return (ev, z) if vectors else ev


def tmatrix(alpha, beta):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    n = len(a)
    eig = np.zeros(n)
    err = np.zeros(n)
    mult = np.zeros(n, dtype=np.int32)
    nout = ctypes.c_int(0)
    th = ctypes.c_double(0)
    mt = ctypes.c_double(0)
    rc = lib().oracle_tmatrix(n, _p(a), _p(b), _p(eig), _p(err), _p(mult), ctypes.byref(nout), ctypes.byref(th), ctypes.byref(mt))
    assert rc == 0
    k = nout.value
    return dict(eig=eig[:k], err=err[:k], mult=mult[:k], thold=th.value, multol=mt.value)


def start_value(seed, i, lo=-1.0, hi=1.0):
    return lib().oracle_start_value(seed, i, lo, hi)


def ground_and_gap(sectors):
    """Ground-state energy and gap over per-sector eigenvalue lists."""
    counts = np.array([len(v) for v in sectors], dtype=np.int32)
    vals = np.concatenate([np.asarray(v, dtype=np.float64) for v in sectors] + [np.zeros(0)])
    g, d = ctypes.c_double(), ctypes.c_double()
    assert lib().oracle_ground_and_gap(len(sectors), _p(counts), _p(vals), ctypes.byref(g), ctypes.byref(d)) == 0
    return g.value, d.value


def sparsediag(m: Model, n_up, max_iter=0, nev=1, seed=0, fixed=False, max_ab=4096):
    ev = np.zeros(max(nev, 1))
    a = np.zeros(max_ab)
    b = np.zeros(max_ab)
    n_ev, n_ab, its, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    dim = ctypes.c_long()
    rc = lib().oracle_sparsediag(m.n_sites, n_up, *m.cargs(), max_iter, nev, seed, int(fixed),
                                 _p(ev), ctypes.byref(n_ev), _p(a), _p(b), max_ab, ctypes.byref(n_ab),
                                 ctypes.byref(its), ctypes.byref(err), ctypes.byref(dim))
    assert rc == 0
    na = min(n_ab.value, max_ab)
    return dict(ev=ev[: n_ev.value].copy(), alpha=a[:na].copy(), beta=b[:na].copy(),
                iterations=its.value, error_code=err.value, dim=dim.value)


def bench_hv(m: Model, n_up, rows, reps=3, threads=1):
    """Median seconds of `reps` map-of-maps H.v over rows [0, rows), on
    `threads` threads, and the build time of those rows."""
    bs = ctypes.c_double()
    nnz = ctypes.c_long()
    t = lib().oracle_bench_hv(m.n_sites, n_up, *m.cargs(), rows, reps, threads, ctypes.byref(bs),
                              ctypes.byref(nnz))
    return dict(seconds=t, build_seconds=bs.value, nnz=nnz.value)


PHASES = ("basis", "matrix", "hv", "vector", "tmatrix")


def sparsediag_timed(m: Model, n_up, max_iter=0, nev=1, seed=0, fixed=False):
    """The oracle's do_subspace with wall seconds per phase."""
    ev = np.zeros(max(nev, 1))
    ph = np.zeros(len(PHASES))
    n_ev, its, dim = ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
    rc = lib().oracle_sparsediag_timed(m.n_sites, n_up, *m.cargs(), max_iter, nev, seed, int(fixed), _p(ev),
                                       ctypes.byref(n_ev), ctypes.byref(its), ctypes.byref(dim), _p(ph))
    assert rc == 0
    return dict(ev=ev[: n_ev.value].copy(), iterations=its.value, dim=dim.value,
                phases=dict(zip(PHASES, ph.tolist())))


SITE_OPS = {"Sz": 0, "Splus": 1, "Sminus": 2, "Sx": 3}


def expect_site_term(N, n_up, v, a, i, b=None, j=-1):
    """<v|O|v> by the oracle's operator-matrix restatement: O = a(i) when
    j < 0, a(i)*b(i) when j == i, a(i)*b(j) otherwise (names of SITE_OPS)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = ctypes.c_double()
    rc = lib().oracle_expect_site_term(N, n_up, _p(v), SITE_OPS[a], i, SITE_OPS[b] if b else 0, j,
                                       ctypes.byref(out))
    assert rc == 0
    return out.value


def translation_momenta(name, L, W=0, H=0):
    out = np.zeros(4096)
    nd = ctypes.c_int()
    n = lib().oracle_translation_momenta(name.encode(), L, W, H, 4096 // 3, _p(out), ctypes.byref(nd))
    assert n >= 0
    return out[: n * nd.value].reshape(n, nd.value)


def translations(name, L, W=0, H=0, k=()):
    """(phases [n_t] complex, perms [n_t, N]) of the lattice at momentum k."""
    N = lattice(name, L, W, H)[0]
    k = np.ascontiguousarray(k, dtype=np.float64)
    perm = np.zeros((1024, N), dtype=np.int64)
    re, im = np.zeros(1024), np.zeros(1024)
    n = lib().oracle_translations(name.encode(), L, W, H, len(k), _p(k), 1024, _p(perm), _p(re), _p(im))
    assert n > 0
    return re[:n] + 1j * im[:n], perm[:n].copy()


def _trans_args(phases, perms):
    perms = np.ascontiguousarray(perms, dtype=np.int64)
    re = np.ascontiguousarray(np.real(phases), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(phases), dtype=np.float64)
    return len(re), perms, re, im


def bloch_basis(N, n_up, phases, perms):
    """Representatives (bit-encoded) and normalizations; n_up < 0 for the
    whole space."""
    nt, perms, re, im = _trans_args(phases, perms)
    cap = 1 << min(N, 20)
    reps = np.zeros(cap, dtype=np.uint32)
    norm = np.zeros(cap)
    n = lib().oracle_bloch_basis(N, n_up, nt, _p(perms), _p(re), _p(im), cap, _p(reps), _p(norm))
    assert n >= 0
    return reps[:n].copy(), norm[:n].copy()


def bloch_dense(m: Model, n_up, phases, perms):
    """Dense complex H of the model in the Bloch basis of n_up up spins."""
    nt, perms, re, im = _trans_args(phases, perms)
    one = np.zeros(1)
    D = -lib().oracle_bloch_dense(m.n_sites, n_up, *m.cargs(), nt, _p(perms), _p(re), _p(im), 0, _p(one), _p(one))
    assert D >= 0, D
    out_re = np.zeros(max(D * D, 1))
    out_im = np.zeros(max(D * D, 1))
    D = lib().oracle_bloch_dense(m.n_sites, n_up, *m.cargs(), nt, _p(perms), _p(re), _p(im), D,
                                 _p(out_re), _p(out_im))
    assert D >= 0, D
    return (out_re[: D * D] + 1j * out_im[: D * D]).reshape(D, D)


def ring24_fixture(n_up, nev):
    """The oracle's FP64 sparsediag of the L=24 Heisenberg ring for one
    (n_up, nev), stored by tests/golden/make_ring24_fixture.py."""
    import json

    with open(os.path.join(ROOT, "tests", "golden", "ring24_oracle.json")) as f:
        runs = json.load(f)["runs"]
    for r in runs:
        if r["n_up"] == n_up and r["nev"] == nev:
            return r
    raise KeyError((n_up, nev))
