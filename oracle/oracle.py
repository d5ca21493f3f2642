"""ctypes front end of the sqaod CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/_build/libsqaod_oracle.so (a plain-C restatement of sqaod.cpu,
see sqaod_oracle.h) for the tests, the smoke check and the CPU-baseline leg
of bench.py.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsqaod_oracle.so")
REF_RANDOM_PATH = os.path.join(_HERE, "_ref", "libsqaod_ref_random.so")

ALGO_NAIVE, ALGO_COLORING, ALGO_SA_NAIVE, ALGO_SA_COLORING = 2, 3, 6, 7
MINIMIZE, MAXIMIZE = 0, 1
STREAM_MT, STREAM_PHILOX = 0, 1

_lib = None


def build():
    """Compile the oracle with its Makefile (and the reference RNG when the
    reference tree is present)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if os.path.exists("/root/reference/sqaodc/common/Random.cpp"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, u64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double
        sz = ctypes.c_size_t
        sigs = {
            "sqo_mt_stream": (None, [u64, u64, vp, sz]),
            "sqo_dot_f32": (ctypes.c_float, [vp, vp, i32]),
            "sqo_dot_f64": (f64, [vp, vp, i32]),
            "sqo_dg_hamiltonian_f32": (None, [i32, vp, vp, vp, vp]),
            "sqo_dg_hamiltonian_f64": (None, [i32, vp, vp, vp, vp]),
            "sqo_dg_energy_q_f64acc": (None, [i32, i32, vp, vp, f64, vp, vp]),
            "sqo_dg_energy_x_f64acc": (None, [i32, i32, vp, vp, vp]),
            "sqo_dg_new": (vp, [i32, i32]),
            "sqo_dg_free": (None, [vp]),
            "sqo_dg_seed": (None, [vp, u64]),
            "sqo_dg_set_qubo": (None, [vp, i32, vp, i32]),
            "sqo_dg_set_hamiltonian": (None, [vp, i32, vp, vp, f64]),
            "sqo_dg_get_hamiltonian": (None, [vp, vp, vp, vp]),
            "sqo_dg_set_m": (None, [vp, i32]),
            "sqo_dg_get_m": (i32, [vp]),
            "sqo_dg_set_algorithm": (None, [vp, i32]),
            "sqo_dg_prepare": (i32, [vp]),
            "sqo_dg_randomize_spin": (None, [vp]),
            "sqo_dg_set_q": (None, [vp, vp]),
            "sqo_dg_get_q": (None, [vp, vp]),
            "sqo_dg_anneal_one_step": (None, [vp, f64, f64]),
            "sqo_dg_calculate_E": (None, [vp, vp]),
            "sqo_dg_get_system_E": (f64, [vp, f64, f64]),
            "sqo_dg_last_flips": (ctypes.c_longlong, [vp]),
            "sqo_bg_new": (vp, [i32, i32]),
            "sqo_bg_free": (None, [vp]),
            "sqo_bg_seed": (None, [vp, u64]),
            "sqo_bg_set_qubo": (None, [vp, i32, i32, vp, vp, vp, i32]),
            "sqo_bg_set_hamiltonian": (None, [vp, i32, i32, vp, vp, vp, f64]),
            "sqo_bg_get_hamiltonian": (None, [vp, vp, vp, vp, vp]),
            "sqo_bg_set_m": (None, [vp, i32]),
            "sqo_bg_get_m": (i32, [vp]),
            "sqo_bg_set_algorithm": (None, [vp, i32]),
            "sqo_bg_prepare": (i32, [vp]),
            "sqo_bg_randomize_spin": (None, [vp]),
            "sqo_bg_set_q": (None, [vp, vp, vp]),
            "sqo_bg_get_q": (None, [vp, vp, vp]),
            "sqo_bg_anneal_one_step": (None, [vp, f64, f64]),
            "sqo_bg_calculate_E": (None, [vp, vp]),
            "sqo_bg_get_system_E": (f64, [vp, f64, f64]),
            "sqo_bg_hamiltonian_f64": (None, [i32, i32, vp, vp, vp, vp, vp, vp, vp]),
            "sqo_bg_energy_q_f64acc": (None, [i32, i32, i32, vp, vp, vp, f64, vp, vp, vp]),
            "sqo_dg_bf_search": (i32, [i32, i32, vp, i32, i32, vp, vp]),
            "sqo_bg_bf_search": (i32, [i32, i32, i32, vp, vp, vp, i32, i32, i32, vp, vp, vp]),
            "sqo_expf_checksum": (u64, [u64, u64]),
            "sqo_philox4x32_10": (None, [vp, vp, vp]),
            "sqo_philox_attempts": (None, [u64, u64, i32, i32, vp, vp, i32]),
            "sqo_philox_spins": (None, [u64, ctypes.c_uint32, i32, i32, vp, sz, i32]),
            "sqo_philox_naive": (None, [u64, u64, u64, vp]),
            "sqo_philox_bg": (None, [u64, u64, i32, u64, vp]),
            "sqo_dg_set_stream": (None, [vp, i32]),
            "sqo_dg_set_local_fields": (None, [vp, i32]),
            "sqo_dg_anneal_one_step_injected": (None, [vp, f64, f64, vp, vp, vp]),
            "sqo_bg_set_stream": (None, [vp, i32]),
            "sqo_bg_set_local_fields": (None, [vp, i32]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dt(dtype):
    return 1 if np.dtype(dtype) == np.float64 else 0


# ---------------- RNG ----------------
def mt_stream(seed, n, skip=0):
    out = np.empty(n, np.uint32)
    lib().sqo_mt_stream(seed, skip, _p(out), n)
    return out


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block (Salmon et al., SC'11)."""
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().sqo_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def philox_attempts(seed, sweep, N, m, dtype=np.float32):
    """The rocm device's attempt arrays (X, U) of one dense sweep."""
    X = np.empty((m, N), np.uint32)
    U = np.empty((m, N), dtype)
    lib().sqo_philox_attempts(seed, sweep, N, m, _p(X), _p(U), _dt(dtype))
    return X, U


def philox_spins(seed, salt, N, m):
    q = np.empty((m, N), np.float64)
    lib().sqo_philox_spins(seed, salt, N, m, _p(q), 8, N)
    return q.astype(np.int8)


def expf_checksum(lo=0, hi=1 << 32):
    return lib().sqo_expf_checksum(lo, hi)


def dot(v0, v1):
    v0 = np.ascontiguousarray(v0)
    v1 = np.ascontiguousarray(v1, dtype=v0.dtype)
    if v0.dtype == np.float32:
        return lib().sqo_dot_f32(_p(v0), _p(v1), len(v0))
    return lib().sqo_dot_f64(_p(v0), _p(v1), len(v0))


# ---------------- dense formulas ----------------
def dense_hamiltonian(W, dtype=np.float64):
    W = np.ascontiguousarray(W, dtype=dtype)
    N = W.shape[0]
    h = np.empty(N, dtype)
    J = np.empty((N, N), dtype)
    c = np.empty(1, dtype)
    fn = lib().sqo_dg_hamiltonian_f64 if _dt(dtype) else lib().sqo_dg_hamiltonian_f32
    fn(N, _p(W), _p(h), _p(J), _p(c))
    return h, J, c[0]


def dense_energy_q(h, J, c, q):
    """FP64 E_y = -c - h.q_y - q_y^T J q_y for q of shape (m, N)."""
    h = np.ascontiguousarray(h, np.float64)
    J = np.ascontiguousarray(J, np.float64)
    q = np.ascontiguousarray(np.atleast_2d(q), np.int8)
    m, N = q.shape
    E = np.empty(m, np.float64)
    lib().sqo_dg_energy_q_f64acc(N, m, _p(h), _p(J), float(c), _p(q), _p(E))
    return E


def dense_energy_x(W, x):
    W = np.ascontiguousarray(W, np.float64)
    x = np.ascontiguousarray(np.atleast_2d(x), np.int8)
    nx, N = x.shape
    E = np.empty(nx, np.float64)
    lib().sqo_dg_energy_x_f64acc(N, nx, _p(W), _p(x), _p(E))
    return E


# ---------------- dense annealer ----------------
class DenseGraphAnnealer:
    """Oracle counterpart of sqaod.cpu.DenseGraphAnnealer (serial coloring when
    n_workers == 1, the OpenMP static-chunk variant otherwise)."""

    def __init__(self, dtype=np.float32, n_workers=1):
        self.dtype = np.dtype(dtype)
        self._h = lib().sqo_dg_new(_dt(dtype), n_workers)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().sqo_dg_free(self._h)
            self._h = None

    def seed(self, s):
        lib().sqo_dg_seed(self._h, s)

    def set_qubo(self, W, optimize=MINIMIZE):
        W = np.ascontiguousarray(W, self.dtype)
        self.N = W.shape[0]
        lib().sqo_dg_set_qubo(self._h, self.N, _p(W), optimize)

    def set_hamiltonian(self, h, J, c):
        h = np.ascontiguousarray(h, self.dtype)
        J = np.ascontiguousarray(J, self.dtype)
        self.N = J.shape[0]
        lib().sqo_dg_set_hamiltonian(self._h, self.N, _p(h), _p(J), float(c))

    def get_hamiltonian(self):
        h = np.empty(self.N, self.dtype)
        J = np.empty((self.N, self.N), self.dtype)
        c = ctypes.c_double()
        lib().sqo_dg_get_hamiltonian(self._h, _p(h), _p(J), ctypes.byref(c))
        return h, J, self.dtype.type(c.value)

    def set_m(self, m):
        lib().sqo_dg_set_m(self._h, m)

    @property
    def m(self):
        return lib().sqo_dg_get_m(self._h)

    def set_algorithm(self, algo):
        lib().sqo_dg_set_algorithm(self._h, algo)

    def prepare(self):
        return lib().sqo_dg_prepare(self._h)

    def randomize_spin(self):
        lib().sqo_dg_randomize_spin(self._h)

    def set_q(self, q):
        q = np.ascontiguousarray(q, np.int8)
        assert q.shape == (self.m, self.N)
        lib().sqo_dg_set_q(self._h, _p(q))

    def get_q(self):
        q = np.empty((self.m, self.N), np.int8)
        lib().sqo_dg_get_q(self._h, _p(q))
        return q

    def anneal_one_step(self, G, beta):
        lib().sqo_dg_anneal_one_step(self._h, float(G), float(beta))

    def calculate_E(self):
        E = np.empty(self.m, self.dtype)
        lib().sqo_dg_calculate_E(self._h, _p(E))
        return E

    def get_system_E(self, G, beta):
        return lib().sqo_dg_get_system_E(self._h, float(G), float(beta))

    def last_flips(self):
        return lib().sqo_dg_last_flips(self._h)

    def set_stream(self, kind):
        """STREAM_MT (sqaod.cpu's per-worker MT19937) or STREAM_PHILOX (the
        rocm device's counter layout)."""
        lib().sqo_dg_set_stream(self._h, kind)

    def set_local_fields(self, on=True):
        """Incremental local fields instead of the per-attempt dot (the same
        results for integer-valued problems)."""
        lib().sqo_dg_set_local_fields(self._h, 1 if on else 0)

    def anneal_one_step_injected(self, G, beta, X, U, Y=None):
        """One step consuming the given stream: X, U of shape (m, N) in (slice,
        attempt) order; for the naive algorithm Y, X, U of N m entries."""
        X = np.ascontiguousarray(X, np.uint32)
        U = np.ascontiguousarray(U, self.dtype)
        Yp = None
        if Y is not None:
            Y = np.ascontiguousarray(Y, np.uint32)
            Yp = _p(Y)
        lib().sqo_dg_anneal_one_step_injected(self._h, float(G), float(beta), Yp, _p(X), _p(U))


class BipartiteGraphAnnealer:
    """Oracle counterpart of sqaod.cpu.BipartiteGraphAnnealer (serial when
    n_workers == 1, the OpenMP static-chunk variant otherwise)."""

    def __init__(self, dtype=np.float32, n_workers=1):
        self.dtype = np.dtype(dtype)
        self._h = lib().sqo_bg_new(_dt(dtype), n_workers)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().sqo_bg_free(self._h)
            self._h = None

    def seed(self, s):
        lib().sqo_bg_seed(self._h, s)

    def set_qubo(self, b0, b1, W, optimize=MINIMIZE):
        b0 = np.ascontiguousarray(b0, self.dtype)
        b1 = np.ascontiguousarray(b1, self.dtype)
        W = np.ascontiguousarray(W, self.dtype)
        self.N0, self.N1 = len(b0), len(b1)
        assert W.shape == (self.N1, self.N0)
        lib().sqo_bg_set_qubo(self._h, self.N0, self.N1, _p(b0), _p(b1), _p(W), optimize)

    def set_hamiltonian(self, h0, h1, J, c):
        h0 = np.ascontiguousarray(h0, self.dtype)
        h1 = np.ascontiguousarray(h1, self.dtype)
        J = np.ascontiguousarray(J, self.dtype)
        self.N0, self.N1 = len(h0), len(h1)
        lib().sqo_bg_set_hamiltonian(self._h, self.N0, self.N1, _p(h0), _p(h1), _p(J), float(c))

    def get_hamiltonian(self):
        h0 = np.empty(self.N0, self.dtype)
        h1 = np.empty(self.N1, self.dtype)
        J = np.empty((self.N1, self.N0), self.dtype)
        c = ctypes.c_double()
        lib().sqo_bg_get_hamiltonian(self._h, _p(h0), _p(h1), _p(J), ctypes.byref(c))
        return h0, h1, J, self.dtype.type(c.value)

    def set_m(self, m):
        lib().sqo_bg_set_m(self._h, m)

    @property
    def m(self):
        return lib().sqo_bg_get_m(self._h)

    def set_algorithm(self, algo):
        lib().sqo_bg_set_algorithm(self._h, algo)

    def prepare(self):
        return lib().sqo_bg_prepare(self._h)

    def randomize_spin(self):
        lib().sqo_bg_randomize_spin(self._h)

    def set_q(self, q0, q1):
        q0 = np.ascontiguousarray(q0, np.int8)
        q1 = np.ascontiguousarray(q1, np.int8)
        lib().sqo_bg_set_q(self._h, _p(q0), _p(q1))

    def get_q(self):
        q0 = np.empty((self.m, self.N0), np.int8)
        q1 = np.empty((self.m, self.N1), np.int8)
        lib().sqo_bg_get_q(self._h, _p(q0), _p(q1))
        return q0, q1

    def anneal_one_step(self, G, beta):
        lib().sqo_bg_anneal_one_step(self._h, float(G), float(beta))

    def calculate_E(self):
        E = np.empty(self.m, self.dtype)
        lib().sqo_bg_calculate_E(self._h, _p(E))
        return E

    def get_system_E(self, G, beta):
        return lib().sqo_bg_get_system_E(self._h, float(G), float(beta))

    def set_stream(self, kind):
        lib().sqo_bg_set_stream(self._h, kind)

    def set_local_fields(self, on=True):
        lib().sqo_bg_set_local_fields(self._h, 1 if on else 0)


def bipartite_hamiltonian(b0, b1, W):
    b0 = np.ascontiguousarray(b0, np.float64)
    b1 = np.ascontiguousarray(b1, np.float64)
    W = np.ascontiguousarray(W, np.float64)
    N0, N1 = len(b0), len(b1)
    h0, h1 = np.empty(N0), np.empty(N1)
    J = np.empty((N1, N0))
    c = np.empty(1)
    lib().sqo_bg_hamiltonian_f64(N0, N1, _p(b0), _p(b1), _p(W), _p(h0), _p(h1), _p(J), _p(c))
    return h0, h1, J, c[0]


def bipartite_energy_q(h0, h1, J, c, q0, q1):
    h0 = np.ascontiguousarray(h0, np.float64)
    h1 = np.ascontiguousarray(h1, np.float64)
    J = np.ascontiguousarray(J, np.float64)
    q0 = np.ascontiguousarray(np.atleast_2d(q0), np.int8)
    q1 = np.ascontiguousarray(np.atleast_2d(q1), np.int8)
    m = q0.shape[0]
    E = np.empty(m)
    lib().sqo_bg_energy_q_f64acc(len(h0), len(h1), m, _p(h0), _p(h1), _p(J), float(c), _p(q0), _p(q1), _p(E))
    return E


def dense_bf_search(W, dtype=np.float32, optimize=MINIMIZE, tile_size=1024):
    """Returns (Emin, xs) with xs the packed minimisers (bit N-1-pos = x[pos]),
    ascending, at most tile_size of them."""
    W = np.ascontiguousarray(W, np.float64)
    N = W.shape[0]
    xs = np.empty(max(1, min(tile_size, 1 << N)), np.uint64)
    Emin = ctypes.c_double()
    n = lib().sqo_dg_bf_search(_dt(dtype), N, _p(W), optimize, tile_size, ctypes.byref(Emin), _p(xs))
    if n < 0:
        raise ValueError("N out of range")
    return Emin.value, xs[:n].copy()


def bipartite_bf_search(b0, b1, W, dtype=np.float32, optimize=MINIMIZE, tile_size0=1024, tile_size1=1024):
    """Returns (Emin, xs0, xs1): the packed minimising pairs in the one-worker
    visiting order, at most N0 + N1 of them."""
    b0 = np.ascontiguousarray(b0, np.float64)
    b1 = np.ascontiguousarray(b1, np.float64)
    W = np.ascontiguousarray(W, np.float64)
    N0, N1 = len(b0), len(b1)
    xs0 = np.empty(N0 + N1, np.uint64)
    xs1 = np.empty(N0 + N1, np.uint64)
    Emin = ctypes.c_double()
    n = lib().sqo_bg_bf_search(_dt(dtype), N0, N1, _p(b0), _p(b1), _p(W), optimize, tile_size0, tile_size1,
                               ctypes.byref(Emin), _p(xs0), _p(xs1))
    if n < 0:
        raise ValueError("N0, N1 out of range")
    return Emin.value, xs0[:n].copy(), xs1[:n].copy()


def unpack_bits(packed, N):
    """Packed x -> int8 bit vector, element pos = bit (N-1-pos) (Common.cpp:88-93)."""
    packed = np.atleast_1d(np.asarray(packed, np.uint64))
    shifts = np.arange(N - 1, -1, -1, dtype=np.uint64)
    return ((packed[:, None] >> shifts[None, :]) & np.uint64(1)).astype(np.int8)


# ---------------- reference MT19937 compiled from the reference tree ----------------
_ref = None


def ref_random():
    """The reference's own sqaod::Random (oracle/_ref), or None if it is not built."""
    global _ref
    if _ref is None and os.path.exists(REF_RANDOM_PATH):
        L = ctypes.CDLL(REF_RANDOM_PATH)
        vp, u64, sz = ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_size_t
        L.sqref_mt_stream.argtypes = [u64, u64, vp, sz]
        L.sqref_mt_f32.argtypes = [u64, vp, sz]
        L.sqref_mt_f64.argtypes = [u64, vp, sz]
        L.sqref_mt_randint.argtypes = [u64, ctypes.c_int, vp, sz]
        for f in (L.sqref_mt_stream, L.sqref_mt_f32, L.sqref_mt_f64, L.sqref_mt_randint):
            f.restype = None
        _ref = L
    return _ref
