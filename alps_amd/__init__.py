"""alps_amd -- MI355X (gfx950) exact-diagonalization engine for the ALPS
`sparsediag` hot path: fixed-Sz spin-1/2 Hamiltonians, matrix-free H.v and
the ietl::lanczos recurrence, behind the C ABI of include/alps_ed.h.

This module is a thin ctypes binding of that ABI (libalps_ed.so, built by
alps_amd/build.py).  There is no fallback: if the library is missing or no
GPU is present, the calls raise.
"""
import ctypes
import json
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libalps_ed.so")

F32, F64 = 0, 1
_DTYPES = {"f32": F32, "float32": F32, F32: F32, "f64": F64, "float64": F64, F64: F64}

_lib = None


class EDError(RuntimeError):
    """Error status returned by the engine (see include/alps_ed.h)."""

    def __init__(self, status, message):
        super().__init__(f"alps_ed status {status}: {message}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"alps_amd: engine library not built: {LIB_PATH} (run python -m alps_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, u64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        sig = {
            "alps_ed_version": ([], i32),
            "alps_ed_last_error": ([], ctypes.c_char_p),
            "alps_ed_device_info": ([ctypes.c_char_p, ctypes.c_size_t, P, P], i32),
            "alps_ed_set_stream": ([P], i32),
            "alps_ed_synchronize": ([], i32),
            "alps_ed_operator_create": ([P, P], i32),
            "alps_ed_operator_create_momentum": ([P, i32, P, P, P, P], i32),
            "alps_ed_sparsediag_momentum": ([P, i32, P, P, P, P, P], i32),
            "alps_ed_operator_destroy": ([P], i32),
            "alps_ed_operator_schedule": ([P, P, P], i32),
            "alps_ed_operator_set_schedule": ([P, i32], i32),
            "alps_ed_operator_dimension": ([P, P], i32),
            "alps_ed_operator_bytes": ([P, P], i32),
            "alps_ed_operator_states": ([P, u64, u64, P], i32),
            "alps_ed_operator_index": ([P, P, u64, P], i32),
            "alps_ed_operator_elements": ([P, u64, u64, i32, P, P, P], i32),
            "alps_ed_vector_create": ([u64, i32, P], i32),
            "alps_ed_vector_destroy": ([P], i32),
            "alps_ed_vector_size": ([P, P], i32),
            "alps_ed_vector_dtype": ([P, P], i32),
            "alps_ed_vector_data": ([P, P], i32),
            "alps_ed_vector_upload": ([P, P, u64], i32),
            "alps_ed_vector_download": ([P, P, u64], i32),
            "alps_ed_vector_fill_random": ([P, u64, f64, f64], i32),
            "alps_ed_vector_fill": ([P, f64], i32),
            "alps_ed_vector_copy": ([P, P], i32),
            "alps_ed_vector_scale": ([P, f64], i32),
            "alps_ed_vector_axpby": ([P, f64, P, f64], i32),
            "alps_ed_vector_dot": ([P, P, P], i32),
            "alps_ed_vector_nrm2": ([P, P], i32),
            "alps_ed_matvec": ([P, P, P], i32),
            "alps_ed_vector_spin_moments": ([P, P, P, P, P], i32),
            "alps_ed_vector_flip_pairs": ([P, P, i32, P, P, P], i32),
            "alps_ed_vector_bloch_expand": ([P, P, P, P], i32),
            "alps_ed_lanczos_step_vectors": ([P, P, P, P, f64, P, P], i32),
            "alps_ed_lanczos_step_vectors_noalloc": ([P, P, P, P, f64, P, P], i32),
            "alps_ed_lanczos_create": ([P, i32, i32, i32, P], i32),
            "alps_ed_lanczos_ritz": ([P, i32, i32, P, P], i32),
            "alps_ed_lanczos_destroy": ([P], i32),
            "alps_ed_lanczos_set_reorthogonalization": ([P, i32], i32),
            "alps_ed_lanczos_reorthogonalizations": ([P, P], i32),
            "alps_ed_lanczos_start": ([P, u64, P], i32),
            "alps_ed_lanczos_steps": ([P, i32], i32),
            "alps_ed_lanczos_coefficients": ([P, P, P, i32, P], i32),
            "alps_ed_lanczos_vectors": ([P, P, P], i32),
            "alps_ed_lanczos_snapshot_size": ([P, P], i32),
            "alps_ed_lanczos_save": ([P, P, ctypes.c_size_t], i32),
            "alps_ed_lanczos_load": ([P, ctypes.c_char_p, ctypes.c_size_t], i32),
            "alps_ed_tmatrix_analyze": ([i32, P, P, P, P, P, P], i32),
            "alps_ed_sparsediag": ([P, P, P], i32),
            "alps_ed_run_parameter_file": ([ctypes.c_char_p, P, ctypes.c_size_t, P], i32),
            "alps_ed_batch_create": ([P, P], i32),
            "alps_ed_batch_destroy": ([P], i32),
            "alps_ed_batch_submit": ([P, P, P, P], i32),
            "alps_ed_batch_submit_momentum": ([P, P, i32, P, P, P, P, P], i32),
            "alps_ed_batch_progress": ([P, P, P], i32),
            "alps_ed_batch_run": ([P], i32),
            "alps_ed_batch_completed": ([P, P, i32, P], i32),
            "alps_ed_batch_result": ([P, u64, P, P], i32),
            "alps_ed_batch_stats": ([P, P], i32),
            "alps_ed_batch_footprint": ([P, P, i32, P], i32),
            "alps_ed_batch_footprint_momentum": ([P, i32, P, P, P, P, i32, P], i32),
            "alps_ed_describe_parameter_file": ([ctypes.c_char_p, P, ctypes.c_size_t, P], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise EDError(status, lib().alps_ed_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class _Model(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int), ("n_up", ctypes.c_int), ("n_bonds", ctypes.c_int),
                ("bond_site1", ctypes.c_void_p), ("bond_site2", ctypes.c_void_p),
                ("jz", ctypes.c_void_p), ("jxy", ctypes.c_void_p), ("k", ctypes.c_void_p),
                ("h", ctypes.c_void_p), ("d", ctypes.c_void_p)]


class _SolveParams(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("nev", ctypes.c_int), ("fixed_iterations", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("dtype", ctypes.c_int)]


class _SolveResult(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_uint64), ("n_eigenvalues", ctypes.c_int),
                ("eigenvalues", ctypes.c_double * 64), ("iterations", ctypes.c_int),
                ("error_code", ctypes.c_int), ("seconds", ctypes.c_double), ("storage", ctypes.c_int),
                ("reorthogonalizations", ctypes.c_int)]


class _BatchConfig(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("hbm_bytes", ctypes.c_uint64), ("host_bytes", ctypes.c_uint64),
                ("steps_per_round", ctypes.c_int), ("max_active", ctypes.c_int), ("host_threads", ctypes.c_int),
                ("graphs", ctypes.c_int)]


class _BatchStatus(ctypes.Structure):
    _fields_ = [("n_active", ctypes.c_int), ("n_pending", ctypes.c_int), ("hbm_used", ctypes.c_uint64),
                ("hbm_bytes", ctypes.c_uint64), ("rounds", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("solves_done", ctypes.c_uint64), ("seconds_admit", ctypes.c_double),
                ("seconds_steps", ctypes.c_double), ("seconds_decide", ctypes.c_double),
                ("seconds_refine", ctypes.c_double), ("hv_rows_f32", ctypes.c_uint64),
                ("hv_rows_f64", ctypes.c_uint64), ("graph_rounds", ctypes.c_uint64),
                ("graph_captures", ctypes.c_uint64)]


_STORAGE = {0: "f32", 1: "f64"}


def _result_dict(r):
    return dict(dim=r.dim, ev=np.array(r.eigenvalues[: r.n_eigenvalues]), iterations=r.iterations,
                error_code=r.error_code, seconds=r.seconds, storage=_STORAGE.get(r.storage),
                reorthogonalizations=r.reorthogonalizations)


@dataclass
class SpinModel:
    """Spin-1/2 model: bonds (site1, site2) in edge order with per-bond Jz,
    Jxy, K and per-site h, D (the "spin" HAMILTONIAN of the ALPS model
    library, lib/xml/models.xml.in:342-373), in the sector with n_up up spins."""
    n_sites: int
    bond_site1: np.ndarray
    bond_site2: np.ndarray
    jz: np.ndarray
    jxy: np.ndarray
    n_up: int = -1
    k: np.ndarray = None
    h: np.ndarray = None
    d: np.ndarray = None

    def __post_init__(self):
        nb = len(self.bond_site1)
        self.bond_site1 = np.ascontiguousarray(self.bond_site1, dtype=np.int32)
        self.bond_site2 = np.ascontiguousarray(self.bond_site2, dtype=np.int32)
        self.jz = np.ascontiguousarray(np.broadcast_to(self.jz, (nb,)), dtype=np.float64)
        self.jxy = np.ascontiguousarray(np.broadcast_to(self.jxy, (nb,)), dtype=np.float64)
        z = 0.0
        self.k = np.ascontiguousarray(np.broadcast_to(z if self.k is None else self.k, (nb,)), dtype=np.float64)
        self.h = np.ascontiguousarray(np.broadcast_to(z if self.h is None else self.h, (self.n_sites,)), dtype=np.float64)
        self.d = np.ascontiguousarray(np.broadcast_to(z if self.d is None else self.d, (self.n_sites,)), dtype=np.float64)
        if self.n_up < 0:
            self.n_up = self.n_sites // 2
        self.cstruct()

    def cstruct(self):
        """The alps_ed_model view of this model (pointers into its arrays),
        kept while the fields are the same objects: building it costs more
        than a batch submit itself."""
        key = (self.n_sites, self.n_up, id(self.bond_site1), id(self.bond_site2), id(self.jz), id(self.jxy),
               id(self.k), id(self.h), id(self.d))
        cached = self.__dict__.get("_cstruct")
        if cached is not None and cached[0] == key:
            return cached[1]
        m = _Model()
        m.n_sites, m.n_up, m.n_bonds = self.n_sites, self.n_up, len(self.bond_site1)
        addr = lambda a: a.__array_interface__["data"][0]
        m.bond_site1, m.bond_site2 = addr(self.bond_site1), addr(self.bond_site2)
        m.jz, m.jxy, m.k, m.h, m.d = addr(self.jz), addr(self.jxy), addr(self.k), addr(self.h), addr(self.d)
        self.__dict__["_cstruct"] = (key, m)
        return m

    def with_sector(self, n_up):
        return SpinModel(self.n_sites, self.bond_site1, self.bond_site2, self.jz, self.jxy, n_up,
                         self.k, self.h, self.d)


def device_info():
    name = ctypes.create_string_buffer(256)
    mem = ctypes.c_uint64()
    ncu = ctypes.c_int()
    _check(lib().alps_ed_device_info(name, 256, ctypes.byref(mem), ctypes.byref(ncu)))
    return dict(name=name.value.decode(), hbm_bytes=mem.value, n_cu=ncu.value)


def synchronize():
    _check(lib().alps_ed_synchronize())


def _translation_args(translations, n_sites):
    """(phases, site maps) as lattice translations list them -> C arrays."""
    phases, perms = translations
    perms = np.ascontiguousarray(perms, dtype=np.int32).reshape(-1, n_sites)
    re = np.ascontiguousarray(np.real(phases), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(phases), dtype=np.float64)
    assert len(re) == perms.shape[0]
    return len(re), perms, re, im


class Operator:
    """Matrix-free device Hamiltonian of one sector (alps_ed_operator).  With
    translations=(phases, site maps) it is the momentum sector of the
    model's Sz sector, whose complex vectors are stored as 2 * sector_dim
    reals, (re, im) per element (the layout of a complex array)."""

    def __init__(self, model: SpinModel, translations=None):
        self.model = model
        self._m = model.cstruct()
        h = ctypes.c_void_p()
        if translations is None:
            _check(lib().alps_ed_operator_create(ctypes.byref(self._m), ctypes.byref(h)))
        else:
            nt, perms, re, im = _translation_args(translations, model.n_sites)
            _check(lib().alps_ed_operator_create_momentum(ctypes.byref(self._m), nt, _p(perms), _p(re), _p(im),
                                                          ctypes.byref(h)))
        self.handle = h
        d = ctypes.c_uint64()
        _check(lib().alps_ed_operator_dimension(self.handle, ctypes.byref(d)))
        self.dim = d.value

    def close(self):
        if getattr(self, "handle", None):
            lib().alps_ed_operator_destroy(self.handle)
            self.handle = None

    __del__ = close

    def states(self, first=0, count=None):
        count = self.dim - first if count is None else count
        out = np.zeros(max(count, 1), dtype=np.uint32)
        _check(lib().alps_ed_operator_states(self.handle, first, count, _p(out)))
        return out[:count]

    def index(self, states):
        s = np.ascontiguousarray(states, dtype=np.uint32)
        out = np.zeros(max(len(s), 1), dtype=np.uint64)
        _check(lib().alps_ed_operator_index(self.handle, _p(s), len(s), _p(out)))
        return out[: len(s)]

    def elements(self, row0=0, nrows=None):
        """Explicit rows as CSR (rowptr, cols, vals) in reference order."""
        nrows = self.dim - row0 if nrows is None else nrows
        mp = len(self.model.bond_site1) + 1
        cols = np.zeros(max(nrows * mp, 1), dtype=np.uint32)
        vals = np.zeros(max(nrows * mp, 1), dtype=np.float64)
        cnt = np.zeros(max(nrows, 1), dtype=np.int32)
        _check(lib().alps_ed_operator_elements(self.handle, row0, nrows, mp, _p(cols), _p(vals), _p(cnt)))
        cnt = cnt[:nrows]
        rowptr = np.zeros(nrows + 1, dtype=np.uint64)
        rowptr[1:] = np.cumsum(cnt)
        sel = (np.arange(mp)[None, :] < cnt[:, None]).reshape(-1)
        return rowptr, cols[: nrows * mp][sel], vals[: nrows * mp][sel]

    def matvec(self, x: "Vector", y: "Vector"):
        _check(lib().alps_ed_matvec(self.handle, x.handle, y.handle))
        return y

    HV_SCHEDULES = 6

    def schedule(self):
        """(chosen H.v schedule for FP32 vectors, seconds per schedule or None
        when the operator was not timed)."""
        s = ctypes.c_int()
        sec = np.zeros(self.HV_SCHEDULES)
        _check(lib().alps_ed_operator_schedule(self.handle, ctypes.byref(s), _p(sec)))
        return s.value, (None if sec[0] < 0 else sec)

    def set_schedule(self, s):
        _check(lib().alps_ed_operator_set_schedule(self.handle, int(s)))

    def spin_moments(self, v: "Vector"):
        """(sum v^2, sum v^2 Sz_i [N], sum v^2 Sz_i Sz_j [N, N]) on the device."""
        N = self.model.n_sites
        n2 = ctypes.c_double()
        sz = np.zeros(N)
        szsz = np.zeros(N * N)
        _check(lib().alps_ed_vector_spin_moments(self.handle, v.handle, ctypes.byref(n2), _p(sz), _p(szsz)))
        return n2.value, sz, szsz.reshape(N, N)

    def flip_pairs(self, v: "Vector", site_i, site_j):
        """<v| S+_i S-_j |v> for each pair (site_i[p], site_j[p])."""
        a = np.ascontiguousarray(site_i, dtype=np.int32)
        b = np.ascontiguousarray(site_j, dtype=np.int32)
        out = np.zeros(max(len(a), 1))
        _check(lib().alps_ed_vector_flip_pairs(self.handle, v.handle, len(a), _p(a), _p(b), _p(out)))
        return out[: len(a)]

    def bloch_expand(self, w: "Vector"):
        """(re, im): a momentum-sector vector in the configuration basis of
        its fixed-Sz sector, as two FP64 device vectors."""
        n = math.comb(self.model.n_sites, self.model.n_up)
        re, im = Vector(n, "f64"), Vector(n, "f64")
        _check(lib().alps_ed_vector_bloch_expand(self.handle, w.handle, re.handle, im.handle))
        return re, im


class Vector:
    """Device vector (alps_ed_vector) stored in FP32 or FP64."""

    def __init__(self, n, dtype="f32"):
        self.n = int(n)
        self.dtype = _DTYPES[dtype]
        h = ctypes.c_void_p()
        _check(lib().alps_ed_vector_create(self.n, self.dtype, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().alps_ed_vector_destroy(self.handle)
            self.handle = None

    __del__ = close

    @classmethod
    def from_numpy(cls, a, dtype="f64"):
        v = cls(len(a), dtype)
        v.upload(a)
        return v

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().alps_ed_vector_upload(self.handle, _p(a), len(a)))

    def numpy(self):
        out = np.zeros(max(self.n, 1))
        _check(lib().alps_ed_vector_download(self.handle, _p(out), self.n))
        return out[: self.n]

    def fill_random(self, seed, lo=-1.0, hi=1.0):
        _check(lib().alps_ed_vector_fill_random(self.handle, seed, lo, hi))

    def fill(self, value):
        _check(lib().alps_ed_vector_fill(self.handle, value))

    def copy_from(self, other):
        _check(lib().alps_ed_vector_copy(self.handle, other.handle))

    def scale(self, a):
        _check(lib().alps_ed_vector_scale(self.handle, a))

    def axpby(self, a, x, b):
        _check(lib().alps_ed_vector_axpby(self.handle, a, x.handle, b))

    def dot(self, other):
        r = ctypes.c_double()
        _check(lib().alps_ed_vector_dot(self.handle, other.handle, ctypes.byref(r)))
        return r.value

    def nrm2(self):
        r = ctypes.c_double()
        _check(lib().alps_ed_vector_nrm2(self.handle, ctypes.byref(r)))
        return r.value

    def data_ptr(self):
        p = ctypes.c_void_p()
        _check(lib().alps_ed_vector_data(self.handle, ctypes.byref(p)))
        return p.value


def lanczos_step_vectors(op, v_prev, v_cur, w, beta_prev, noalloc=False):
    """One make_step on caller vectors; noalloc=True uses the entry point that
    allocates nothing (static device scratch)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    fn = lib().alps_ed_lanczos_step_vectors_noalloc if noalloc else lib().alps_ed_lanczos_step_vectors
    _check(fn(op.handle, v_prev.handle, v_cur.handle, w.handle, beta_prev, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


class Lanczos:
    """Device-resident Lanczos recurrence (alps_ed_lanczos)."""

    def __init__(self, op: Operator, dtype="f32", max_steps=1000, keep_basis=False, reorth=False):
        """reorth=True: partial reorthogonalization (needs keep_basis)."""
        self.op = op
        self.max_steps = max_steps
        h = ctypes.c_void_p()
        _check(lib().alps_ed_lanczos_create(op.handle, _DTYPES[dtype], max_steps, int(keep_basis), ctypes.byref(h)))
        self.handle = h
        if reorth:
            _check(lib().alps_ed_lanczos_set_reorthogonalization(self.handle, 1))

    def reorthogonalizations(self):
        c = ctypes.c_int()
        _check(lib().alps_ed_lanczos_reorthogonalizations(self.handle, ctypes.byref(c)))
        return c.value

    def close(self):
        if getattr(self, "handle", None):
            lib().alps_ed_lanczos_destroy(self.handle)
            self.handle = None

    __del__ = close

    def start(self, seed=0, start: Vector = None):
        _check(lib().alps_ed_lanczos_start(self.handle, seed, None if start is None else start.handle))

    def steps(self, n):
        _check(lib().alps_ed_lanczos_steps(self.handle, n))

    def coefficients(self):
        a = np.zeros(self.max_steps + 1)
        b = np.zeros(self.max_steps + 1)
        n = ctypes.c_int()
        _check(lib().alps_ed_lanczos_coefficients(self.handle, _p(a), _p(b), self.max_steps, ctypes.byref(n)))
        return a[: n.value].copy(), b[: n.value].copy()

    def ritz(self, coeffs):
        """Ritz vectors sum_j coeffs[i, j] v_j over the kept basis."""
        c = np.ascontiguousarray(np.atleast_2d(coeffs), dtype=np.float64)
        nvec, m = c.shape
        outs = [Vector(self.op.dim, "f64") for _ in range(nvec)]
        arr = (ctypes.c_void_p * max(nvec, 1))(*[o.handle for o in outs])
        _check(lib().alps_ed_lanczos_ritz(self.handle, nvec, m, _p(c), arr))
        return outs

    def save(self):
        """Snapshot of the recurrence (alps_ed_lanczos_save) as bytes."""
        n = ctypes.c_size_t()
        _check(lib().alps_ed_lanczos_snapshot_size(self.handle, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(1, n.value))
        _check(lib().alps_ed_lanczos_save(self.handle, buf, n.value))
        return buf.raw[: n.value]

    def load(self, data: bytes):
        """Continues a recurrence from a snapshot of save() (a fresh object
        of the same operator model, dtype, max_steps, keep_basis, reorth)."""
        _check(lib().alps_ed_lanczos_load(self.handle, data, len(data)))

    def vectors(self, dtype="f64"):
        vp, vc = Vector(self.op.dim, dtype), Vector(self.op.dim, dtype)
        _check(lib().alps_ed_lanczos_vectors(self.handle, vp.handle, vc.handle))
        return vp, vc


def tmatrix_analyze(alpha, beta):
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    n = len(a)
    ev, err = np.zeros(n), np.zeros(n)
    mult = np.zeros(n, dtype=np.int32)
    k = ctypes.c_int()
    _check(lib().alps_ed_tmatrix_analyze(n, _p(a), _p(b), _p(ev), _p(err), _p(mult), ctypes.byref(k)))
    return dict(eig=ev[: k.value], err=err[: k.value], mult=mult[: k.value])


def sparsediag(model: SpinModel, max_iter=0, nev=1, seed=0, dtype="f64", fixed=False, translations=None):
    """One sector of SparseDiagMatrix::do_subspace on the device (the
    momentum sector when translations=(phases, site maps) is given)."""
    m = model.cstruct()
    p = _SolveParams(max_iter, nev, int(fixed), seed, _DTYPES[dtype])
    r = _SolveResult()
    if translations is None:
        _check(lib().alps_ed_sparsediag(ctypes.byref(m), ctypes.byref(p), ctypes.byref(r)))
    else:
        nt, perms, re, im = _translation_args(translations, model.n_sites)
        _check(lib().alps_ed_sparsediag_momentum(ctypes.byref(m), nt, _p(perms), _p(re), _p(im), ctypes.byref(p),
                                                 ctypes.byref(r)))
    return _result_dict(r)


def _json_call(fn, text):
    need = ctypes.c_size_t()
    buf = ctypes.create_string_buffer(1 << 16)
    _check(fn(text.encode(), buf, len(buf), ctypes.byref(need)))
    if need.value > len(buf):
        buf = ctypes.create_string_buffer(need.value)
        _check(fn(text.encode(), buf, len(buf), ctypes.byref(need)))
    return json.loads(buf.value.decode())


def describe_parameter_file(text):
    """Tasks of a parameter file with their lattice, couplings and sectors
    resolved on the host (no device needed)."""
    return _json_call(lib().alps_ed_describe_parameter_file, text)["tasks"]


def run_parameter_file(text):
    """The sparsediag task for every task of a parameter file: per-sector
    eigenvalues, ground-state energy and gap."""
    return _json_call(lib().alps_ed_run_parameter_file, text)["tasks"]


class Batch:
    """Many solves advanced together (alps_ed_batch_*): submit requests,
    call progress() (one round) or run() (until all are done), read results
    by ticket."""

    def __init__(self, dtype="f32", hbm_bytes=0, host_bytes=0, steps_per_round=0, max_active=0, host_threads=0,
                 graphs=False):
        cfg = _BatchConfig(_DTYPES[dtype], hbm_bytes, host_bytes, steps_per_round, max_active, host_threads,
                           int(bool(graphs)))
        h = ctypes.c_void_p()
        _check(lib().alps_ed_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().alps_ed_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, model: SpinModel, max_iter=0, nev=1, seed=0, fixed=False, translations=None):
        """Queues one solve (in the momentum sector of translations=(phases,
        site maps) when given); returns its ticket."""
        m = model.cstruct()
        p = _SolveParams(max_iter, nev, int(fixed), seed, F64)
        t = ctypes.c_uint64()
        if translations is None:
            _check(lib().alps_ed_batch_submit(self.handle, ctypes.byref(m), ctypes.byref(p), ctypes.byref(t)))
        else:
            nt, perms, re, im = _translation_args(translations, model.n_sites)
            _check(lib().alps_ed_batch_submit_momentum(self.handle, ctypes.byref(m), nt, _p(perms), _p(re), _p(im),
                                                       ctypes.byref(p), ctypes.byref(t)))
        return t.value

    def progress(self):
        a, p = ctypes.c_int(), ctypes.c_int()
        _check(lib().alps_ed_batch_progress(self.handle, ctypes.byref(a), ctypes.byref(p)))
        return a.value, p.value

    def run(self):
        _check(lib().alps_ed_batch_run(self.handle))

    def completed(self, max_n=65536):
        buf = (ctypes.c_uint64 * max_n)()
        n = ctypes.c_int()
        _check(lib().alps_ed_batch_completed(self.handle, buf, max_n, ctypes.byref(n)))
        return list(buf[: n.value])

    def result(self, ticket):
        r = _SolveResult()
        ready = ctypes.c_int()
        _check(lib().alps_ed_batch_result(self.handle, ticket, ctypes.byref(r), ctypes.byref(ready)))
        return _result_dict(r) if ready.value else None

    def stats(self):
        s = _BatchStatus()
        _check(lib().alps_ed_batch_stats(self.handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _BatchStatus._fields_}


def batch_footprint(model: SpinModel, dtype="f32", max_iter=0, nev=1, translations=None):
    """HBM bytes a batched request would hold while admitted (in the
    momentum sector of translations when given)."""
    m = model.cstruct()
    p = _SolveParams(max_iter, nev, 0, 0, F64)
    b = ctypes.c_uint64()
    if translations is None:
        _check(lib().alps_ed_batch_footprint(ctypes.byref(m), ctypes.byref(p), _DTYPES[dtype], ctypes.byref(b)))
    else:
        nt, perms, re, im = _translation_args(translations, model.n_sites)
        _check(lib().alps_ed_batch_footprint_momentum(ctypes.byref(m), nt, _p(perms), _p(re), _p(im),
                                                      ctypes.byref(p), _DTYPES[dtype], ctypes.byref(b)))
    return b.value
