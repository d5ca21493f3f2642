"""ctypes binding of libsqaod_rocm.so (include/sqaod_rocm.h).

Exposes per-solver namespaces whose function names and argument tuples are
those of the reference's CPython extension modules (the method table of
sqaodc/pyglue/annealer.inc:885-908), so the solver classes call
``cext.anneal_one_step(cobj, G, beta, dtype)`` exactly as the reference's
``DenseGraphAnnealerBase`` does.  The library is required: importing this
module fails loudly when it has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "_lib", "libsqaod_rocm.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libsqaod_rocm.so not found at {LIB_PATH}; build it with `python -m sqaod_amd.build`")

_lib = ctypes.CDLL(LIB_PATH)

_vp, _i32, _u64, _f64, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)

_SIGS = {
    "sqaod_rocm_last_error": (ctypes.c_char_p, []),
    "sqaodc_rocm_version": (None, [ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "sqaod_rocm_device_count": (_i32, [ctypes.POINTER(_i32)]),
    "sqaod_rocm_device_new": (_i32, [_i32, _PP]),
    "sqaod_rocm_device_delete": (_i32, [_vp]),
    "sqaod_rocm_device_synchronize": (_i32, [_vp]),
    "sqaod_rocm_dg_annealer_new": (_i32, [_i32, _PP]),
    "sqaod_rocm_dg_annealer_delete": (_i32, [_vp]),
    "sqaod_rocm_dg_annealer_assign_device": (_i32, [_vp, _vp]),
    "sqaod_rocm_dg_annealer_seed": (_i32, [_vp, _u64]),
    "sqaod_rocm_dg_annealer_set_qubo": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_annealer_set_hamiltonian": (_i32, [_vp, _vp, _vp, _vp, _i32]),
    "sqaod_rocm_dg_annealer_get_problem_size": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_dg_annealer_set_preference": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "sqaod_rocm_dg_annealer_get_preferences": (_i32, [_vp, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]),
    "sqaod_rocm_dg_annealer_get_n_trotters": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_dg_annealer_get_E": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_dg_annealer_get_x": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_annealer_get_hamiltonian": (_i32, [_vp, _vp, _vp, _vp]),
    "sqaod_rocm_dg_annealer_get_q": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_annealer_set_q": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_dg_annealer_set_qset": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_annealer_randomize_spin": (_i32, [_vp]),
    "sqaod_rocm_dg_annealer_calculate_E": (_i32, [_vp]),
    "sqaod_rocm_dg_annealer_prepare": (_i32, [_vp]),
    "sqaod_rocm_dg_annealer_make_solution": (_i32, [_vp]),
    "sqaod_rocm_dg_annealer_get_system_E": (_i32, [_vp, _f64, _f64, ctypes.POINTER(_f64)]),
    "sqaod_rocm_dg_annealer_anneal_one_step": (_i32, [_vp, _f64, _f64]),
    "sqaod_rocm_dg_annealer_get_counters": (_i32, [_vp, ctypes.POINTER(ctypes.c_ulonglong)]),
    "sqaod_rocm_dg_annealer_get_energy_rounding_error": (_i32, [_vp, ctypes.POINTER(_f64)]),
    "sqaod_rocm_dg_annealer_get_E_fp64": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_dg_annealer_get_form_energy_error": (_i32, [_vp, ctypes.POINTER(_f64)]),
    "sqaod_rocm_dg_annealer_set_system_E_trace": (_i32, [_vp, _i32]),
    "sqaod_rocm_dg_annealer_get_system_E_trace": (_i32, [_vp, _vp, _i32, ctypes.POINTER(_i32),
                                                         ctypes.POINTER(ctypes.c_longlong)]),
    "sqaod_rocm_dg_annealer_get_fields": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_annealer_get_phase_times": (_i32, [_vp, ctypes.POINTER(_f64), ctypes.POINTER(ctypes.c_ulonglong)]),
    "sqaod_rocm_dg_bf_searcher_new": (_i32, [_i32, _PP]),
    "sqaod_rocm_dg_bf_searcher_delete": (_i32, [_vp]),
    "sqaod_rocm_dg_bf_searcher_assign_device": (_i32, [_vp, _vp]),
    "sqaod_rocm_dg_bf_searcher_set_qubo": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_bf_searcher_get_problem_size": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_dg_bf_searcher_set_preference": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "sqaod_rocm_dg_bf_searcher_get_preferences": (_i32, [_vp, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]),
    "sqaod_rocm_dg_bf_searcher_prepare": (_i32, [_vp]),
    "sqaod_rocm_dg_bf_searcher_calculate_E": (_i32, [_vp]),
    "sqaod_rocm_dg_bf_searcher_make_solution": (_i32, [_vp]),
    "sqaod_rocm_dg_bf_searcher_search_range": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_u64)]),
    "sqaod_rocm_dg_bf_searcher_search": (_i32, [_vp]),
    "sqaod_rocm_dense_bf_search_noalloc": (_i32, [_i32, _vp, _i32, _i32, _u64, _vp, _vp, _vp, _vp]),
    "sqaod_rocm_dg_bf_searcher_get_n_solutions": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_dg_bf_searcher_get_x": (_i32, [_vp, _vp, _i32, _i32]),
    "sqaod_rocm_dg_bf_searcher_get_E": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_bg_annealer_new": (_i32, [_i32, _PP]),
    "sqaod_rocm_bg_annealer_delete": (_i32, [_vp]),
    "sqaod_rocm_bg_annealer_assign_device": (_i32, [_vp, _vp]),
    "sqaod_rocm_bg_annealer_seed": (_i32, [_vp, _u64]),
    "sqaod_rocm_bg_annealer_set_qubo": (_i32, [_vp, _vp, _i32, _vp, _i32, _vp, _i32]),
    "sqaod_rocm_bg_annealer_set_hamiltonian": (_i32, [_vp, _vp, _i32, _vp, _i32, _vp, _vp]),
    "sqaod_rocm_bg_annealer_get_problem_size": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "sqaod_rocm_bg_annealer_set_preference": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "sqaod_rocm_bg_annealer_get_preferences": (_i32, [_vp, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]),
    "sqaod_rocm_bg_annealer_get_n_trotters": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_bg_annealer_get_E": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_bg_annealer_get_E_fp64": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_bg_annealer_get_x": (_i32, [_vp, _vp, _vp, _i32, _i32, _i32]),
    "sqaod_rocm_bg_annealer_get_hamiltonian": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "sqaod_rocm_bg_annealer_get_q": (_i32, [_vp, _vp, _vp, _i32, _i32, _i32]),
    "sqaod_rocm_bg_annealer_set_q": (_i32, [_vp, _vp, _i32, _vp, _i32]),
    "sqaod_rocm_bg_annealer_set_qset": (_i32, [_vp, _vp, _vp, _i32, _i32, _i32]),
    "sqaod_rocm_bg_annealer_randomize_spin": (_i32, [_vp]),
    "sqaod_rocm_bg_annealer_calculate_E": (_i32, [_vp]),
    "sqaod_rocm_bg_annealer_prepare": (_i32, [_vp]),
    "sqaod_rocm_bg_annealer_make_solution": (_i32, [_vp]),
    "sqaod_rocm_bg_annealer_get_system_E": (_i32, [_vp, _f64, _f64, ctypes.POINTER(_f64)]),
    "sqaod_rocm_bg_annealer_anneal_one_step": (_i32, [_vp, _f64, _f64]),
    "sqaod_rocm_formulas_new": (_i32, [_PP]),
    "sqaod_rocm_formulas_delete": (_i32, [_vp]),
    "sqaod_rocm_formulas_assign_device": (_i32, [_vp, _vp]),
    "sqaod_rocm_dense_graph_batch_calculate_E": (_i32, [_vp, _i32, _vp, _vp, _i32, _vp, _i32]),
    "sqaod_rocm_dense_graph_calculate_hamiltonian": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _i32]),
    "sqaod_rocm_dense_graph_batch_calculate_E_from_spin": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _i32]),
    "sqaod_rocm_bipartite_graph_batch_calculate_E": (_i32, [_vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp,
                                                            _i32]),
    "sqaod_rocm_bipartite_graph_batch_calculate_E_2d": (_i32, [_vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _i32,
                                                               _vp, _i32]),
    "sqaod_rocm_bipartite_graph_calculate_hamiltonian": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32,
                                                                _vp]),
    "sqaod_rocm_bipartite_graph_batch_calculate_E_from_spin": (_i32, [_vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp,
                                                                      _vp, _vp, _i32]),
    "sqaod_rocm_expf_checksum": (_i32, [_u64, _u64, ctypes.POINTER(_u64)]),
    "sqaod_rocm_fp8_decode_table": (_i32, [_vp]),
    "sqaod_rocm_bg_bf_searcher_new": (_i32, [_i32, _PP]),
    "sqaod_rocm_bg_bf_searcher_delete": (_i32, [_vp]),
    "sqaod_rocm_bg_bf_searcher_assign_device": (_i32, [_vp, _vp]),
    "sqaod_rocm_bg_bf_searcher_set_qubo": (_i32, [_vp, _vp, _i32, _vp, _i32, _vp, _i32]),
    "sqaod_rocm_bg_bf_searcher_get_problem_size": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "sqaod_rocm_bg_bf_searcher_set_preference": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "sqaod_rocm_bg_bf_searcher_get_preferences": (_i32, [_vp, ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]),
    "sqaod_rocm_bg_bf_searcher_prepare": (_i32, [_vp]),
    "sqaod_rocm_bg_bf_searcher_calculate_E": (_i32, [_vp]),
    "sqaod_rocm_bg_bf_searcher_make_solution": (_i32, [_vp]),
    "sqaod_rocm_bg_bf_searcher_search_range": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_u64),
                                                      ctypes.POINTER(_u64)]),
    "sqaod_rocm_bg_bf_searcher_search": (_i32, [_vp]),
    "sqaod_rocm_bg_bf_searcher_get_n_solutions": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_bg_bf_searcher_get_x": (_i32, [_vp, _vp, _vp, _i32, _i32, _i32]),
    "sqaod_rocm_bg_bf_searcher_get_E": (_i32, [_vp, _vp, _i32]),
    "sqaod_rocm_dense_batch_new": (_i32, [_PP]),
    "sqaod_rocm_dense_batch_delete": (_i32, [_vp]),
    "sqaod_rocm_dense_batch_set_preference": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "sqaod_rocm_dense_batch_add_qubo": (_i32, [_vp, _vp, _i32, _i32, _u64, _i32, ctypes.POINTER(_i32)]),
    "sqaod_rocm_dense_batch_size": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_dense_batch_prepare": (_i32, [_vp]),
    "sqaod_rocm_dense_batch_anneal": (_i32, [_vp, _vp, _vp, _i32, _i32, ctypes.POINTER(_f64)]),
    "sqaod_rocm_dense_batch_autotune": (_i32, [_vp, _f64, _f64, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "sqaod_rocm_dense_batch_attempts_per_step": (_i32, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "sqaod_rocm_dense_batch_get_shape": (_i32, [_vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "sqaod_rocm_dense_batch_get_E": (_i32, [_vp, _i32, _vp, _i32]),
    "sqaod_rocm_dense_batch_get_x": (_i32, [_vp, _i32, _vp, _i32, _i32]),
    "sqaod_rocm_dense_batch_get_q": (_i32, [_vp, _i32, _vp, _i32, _i32]),
    "sqaod_rocm_server_new": (_i32, [_PP]),
    "sqaod_rocm_server_delete": (_i32, [_vp]),
    "sqaod_rocm_server_set_preference": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "sqaod_rocm_server_submit_dense": (_i32, [_vp, _vp, _i32, _i32, _u64, _vp, _vp, _i32, ctypes.POINTER(_i32)]),
    "sqaod_rocm_server_submit_bipartite": (_i32, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _u64, _vp, _vp, _i32,
                                                  ctypes.POINTER(_i32)]),
    "sqaod_rocm_server_tick": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "sqaod_rocm_server_status": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_u64),
                                        ctypes.POINTER(_u64)]),
    "sqaod_rocm_server_take_done": (_i32, [_vp, _vp, _i32, ctypes.POINTER(_i32)]),
    "sqaod_rocm_server_result_shape": (_i32, [_vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                              ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "sqaod_rocm_server_result": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_longlong)]),
    "sqaod_rocm_server_release": (_i32, [_vp, _i32]),
    "sqaod_rocm_server_result_error": (_i32, [_vp, _i32, ctypes.c_char_p, _sz, ctypes.POINTER(_i32)]),
    "sqaod_rocm_server_now": (_i32, [_vp, ctypes.POINTER(_f64)]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def _check(rc):
    if rc != 0:
        raise RuntimeError(_lib.sqaod_rocm_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_vp)


def _dt(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return 1
    if dt == np.float32:
        return 0
    raise RuntimeError("dtype must be np.float32 or np.float64")


_INT_PREFS = ('n_trotters', 'tile_size', 'slices_per_block', 'refresh_interval', 'profile', 'n_workers',
              'spin_limit')


def _set_prefs(setter, cobj, prefs):
    for k, v in (prefs or {}).items():
        if hasattr(v, '__int__') and not isinstance(v, (str, float)):
            v = int(v)
        _check(setter(cobj, str(k).encode(), str(v).encode()))


def _get_prefs(getter, cobj):
    need = _sz()
    _check(getter(cobj, None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    _check(getter(cobj, buf, need.value, ctypes.byref(need)))
    out = {}
    for item in buf.value.decode().split(';'):
        if not item:
            continue
        k, v = item.split('=', 1)
        out[k] = int(v) if k in _INT_PREFS else v
    return out


def version():
    """(sqaodc API version, HIP version), as sqaodc_cuda_version reports
    (sqaodc version, CUDA version)."""
    a, b = _i32(), _i32()
    _lib.sqaodc_rocm_version(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def device_count():
    n = _i32()
    _check(_lib.sqaod_rocm_device_count(ctypes.byref(n)))
    return n.value


def expf_checksum(lo=0, hi=1 << 32):
    """Device-side checksum of the acceptance-test exp over float bit patterns [lo, hi)."""
    out = _u64()
    _check(_lib.sqaod_rocm_expf_checksum(lo, hi, ctypes.byref(out)))
    return out.value


def fp8_decode_table():
    """The fp8 sweep's device decode of the e4m3 byte codes 0..255 (float32[256])."""
    out = np.empty(256, np.float32)
    _check(_lib.sqaod_rocm_fp8_decode_table(out.ctypes.data_as(ctypes.c_void_p)))
    return out


class _Device:
    """rocm_device module (cuda_device.cpp counterpart)."""

    @staticmethod
    def new(devno=0):
        out = _vp()
        _check(_lib.sqaod_rocm_device_new(devno, ctypes.byref(out)))
        return out.value

    @staticmethod
    def delete(cobj):
        _check(_lib.sqaod_rocm_device_delete(cobj))

    @staticmethod
    def synchronize(cobj):
        _check(_lib.sqaod_rocm_device_synchronize(cobj))


class _DenseGraphAnnealer:
    """rocm_dg_annealer module: annealer.inc:885-908 method table."""

    @staticmethod
    def new(dtype):
        out = _vp()
        _check(_lib.sqaod_rocm_dg_annealer_new(_dt(dtype), ctypes.byref(out)))
        return out.value

    @staticmethod
    def delete(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_delete(cobj))

    @staticmethod
    def assign_device(cobj, devobj, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_assign_device(cobj, devobj))

    @staticmethod
    def seed(cobj, seed, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_seed(cobj, int(seed) & 0xffffffffffffffff))

    @staticmethod
    def set_qubo(cobj, W, optimize, dtype):
        W = np.ascontiguousarray(W, dtype)
        if W.ndim != 2 or W.shape[0] != W.shape[1]:
            raise RuntimeError("W must be a square matrix.")
        _check(_lib.sqaod_rocm_dg_annealer_set_qubo(cobj, _ptr(W), W.shape[0], int(optimize)))

    @staticmethod
    def set_hamiltonian(cobj, h, J, c, dtype):
        h = np.ascontiguousarray(h, dtype)
        J = np.ascontiguousarray(J, dtype)
        cc = np.array([c], dtype)
        if J.ndim != 2 or J.shape[0] != J.shape[1] or h.shape != (J.shape[0],):
            raise RuntimeError("shape mismatch between h and J.")
        _check(_lib.sqaod_rocm_dg_annealer_set_hamiltonian(cobj, _ptr(h), _ptr(J), _ptr(cc), J.shape[0]))

    @staticmethod
    def get_problem_size(cobj, dtype):
        n = _i32()
        _check(_lib.sqaod_rocm_dg_annealer_get_problem_size(cobj, ctypes.byref(n)))
        return n.value

    @staticmethod
    def set_preferences(cobj, prefs, dtype):
        _set_prefs(_lib.sqaod_rocm_dg_annealer_set_preference, cobj, prefs)

    @staticmethod
    def get_preferences(cobj, dtype):
        return _get_prefs(_lib.sqaod_rocm_dg_annealer_get_preferences, cobj)

    @staticmethod
    def _m(cobj):
        m = _i32()
        _check(_lib.sqaod_rocm_dg_annealer_get_n_trotters(cobj, ctypes.byref(m)))
        return m.value

    @staticmethod
    def get_E(cobj, dtype):
        m = _DenseGraphAnnealer._m(cobj)
        E = np.empty(m, dtype)
        _check(_lib.sqaod_rocm_dg_annealer_get_E(cobj, _ptr(E), m))
        return E

    @staticmethod
    def get_x(cobj, dtype):
        m, N = _DenseGraphAnnealer._m(cobj), _DenseGraphAnnealer.get_problem_size(cobj, dtype)
        x = np.empty((m, N), np.int8)
        _check(_lib.sqaod_rocm_dg_annealer_get_x(cobj, _ptr(x), m, N))
        return [row.copy() for row in x]

    @staticmethod
    def get_hamiltonian(cobj, h, J, c, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_get_hamiltonian(cobj, _ptr(h), _ptr(J), _ptr(c)))

    @staticmethod
    def get_q(cobj, dtype):
        m, N = _DenseGraphAnnealer._m(cobj), _DenseGraphAnnealer.get_problem_size(cobj, dtype)
        q = np.empty((m, N), np.int8)
        _check(_lib.sqaod_rocm_dg_annealer_get_q(cobj, _ptr(q), m, N))
        return [row.copy() for row in q]

    @staticmethod
    def set_q(cobj, q, dtype):
        q = np.ascontiguousarray(q, np.int8).reshape(-1)
        _check(_lib.sqaod_rocm_dg_annealer_set_q(cobj, _ptr(q), q.shape[0]))

    @staticmethod
    def set_qset(cobj, qlist, dtype):
        q = np.ascontiguousarray(np.stack([np.asarray(v, np.int8).reshape(-1) for v in qlist]))
        _check(_lib.sqaod_rocm_dg_annealer_set_qset(cobj, _ptr(q), q.shape[0], q.shape[1]))

    @staticmethod
    def randomize_spin(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_randomize_spin(cobj))

    @staticmethod
    def calculate_E(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_calculate_E(cobj))

    @staticmethod
    def prepare(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_prepare(cobj))

    @staticmethod
    def make_solution(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_make_solution(cobj))

    @staticmethod
    def get_system_E(cobj, G, beta, dtype):
        E = _f64()
        _check(_lib.sqaod_rocm_dg_annealer_get_system_E(cobj, float(G), float(beta), ctypes.byref(E)))
        return np.dtype(dtype).type(E.value)

    @staticmethod
    def anneal_one_step(cobj, G, beta, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_anneal_one_step(cobj, float(G), float(beta)))

    @staticmethod
    def get_counters(cobj, dtype):
        c = (ctypes.c_ulonglong * 3)()
        _check(_lib.sqaod_rocm_dg_annealer_get_counters(cobj, c))
        return {'flips': c[0], 'batches': c[1], 'blocked': c[2]}

    @staticmethod
    def get_fields(cobj, dtype):
        m, N = _DenseGraphAnnealer._m(cobj), _DenseGraphAnnealer.get_problem_size(cobj, dtype)
        F = np.empty((m, N), dtype)
        _check(_lib.sqaod_rocm_dg_annealer_get_fields(cobj, _ptr(F), m, N))
        return F

    @staticmethod
    def get_phase_times(cobj, dtype):
        ms = (_f64 * 3)()
        cnt = (ctypes.c_ulonglong * 3)()
        _check(_lib.sqaod_rocm_dg_annealer_get_phase_times(cobj, ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(("refresh", "rng", "sweep"))}

    @staticmethod
    def set_system_E_trace(cobj, capacity, dtype):
        _check(_lib.sqaod_rocm_dg_annealer_set_system_E_trace(cobj, int(capacity)))

    @staticmethod
    def get_system_E_trace(cobj, cap, dtype):
        out = np.empty(max(int(cap), 1), np.float64)
        n, steps = _i32(), ctypes.c_longlong()
        _check(_lib.sqaod_rocm_dg_annealer_get_system_E_trace(cobj, _ptr(out), int(cap), ctypes.byref(n),
                                                              ctypes.byref(steps)))
        return out[:n.value], steps.value

    @staticmethod
    def get_E_fp64(cobj, dtype):
        m = _i32()
        _check(_lib.sqaod_rocm_dg_annealer_get_n_trotters(cobj, ctypes.byref(m)))
        E = np.empty(m.value, np.float64)
        _check(_lib.sqaod_rocm_dg_annealer_get_E_fp64(cobj, _ptr(E), m.value))
        return E

    @staticmethod
    def get_energy_rounding_error(cobj, dtype):
        e = _f64()
        _check(_lib.sqaod_rocm_dg_annealer_get_energy_rounding_error(cobj, ctypes.byref(e)))
        return e.value

    @staticmethod
    def get_form_energy_error(cobj, dtype):
        e = _f64()
        _check(_lib.sqaod_rocm_dg_annealer_get_form_energy_error(cobj, ctypes.byref(e)))
        return e.value


class _DenseGraphBFSearcher:
    """rocm_dg_bf_searcher module: bf_searcher.inc:479-494 method table."""

    @staticmethod
    def new(dtype):
        out = _vp()
        _check(_lib.sqaod_rocm_dg_bf_searcher_new(_dt(dtype), ctypes.byref(out)))
        return out.value

    @staticmethod
    def delete(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_bf_searcher_delete(cobj))

    @staticmethod
    def assign_device(cobj, devobj, dtype):
        _check(_lib.sqaod_rocm_dg_bf_searcher_assign_device(cobj, devobj))

    @staticmethod
    def set_qubo(cobj, W, optimize, dtype):
        W = np.ascontiguousarray(W, dtype)
        if W.ndim != 2 or W.shape[0] != W.shape[1]:
            raise RuntimeError("W must be a square matrix.")
        _check(_lib.sqaod_rocm_dg_bf_searcher_set_qubo(cobj, _ptr(W), W.shape[0], int(optimize)))

    @staticmethod
    def get_problem_size(cobj, dtype):
        n = _i32()
        _check(_lib.sqaod_rocm_dg_bf_searcher_get_problem_size(cobj, ctypes.byref(n)))
        return n.value

    @staticmethod
    def set_preferences(cobj, prefs, dtype):
        _set_prefs(_lib.sqaod_rocm_dg_bf_searcher_set_preference, cobj, prefs)

    @staticmethod
    def get_preferences(cobj, dtype):
        return _get_prefs(_lib.sqaod_rocm_dg_bf_searcher_get_preferences, cobj)

    @staticmethod
    def get_x(cobj, dtype):
        n = _i32()
        _check(_lib.sqaod_rocm_dg_bf_searcher_get_n_solutions(cobj, ctypes.byref(n)))
        N = _DenseGraphBFSearcher.get_problem_size(cobj, dtype)
        x = np.empty((n.value, N), np.int8)
        _check(_lib.sqaod_rocm_dg_bf_searcher_get_x(cobj, _ptr(x), n.value, N))
        return [row.copy() for row in x]

    @staticmethod
    def get_E(cobj, dtype):
        n = _i32()
        _check(_lib.sqaod_rocm_dg_bf_searcher_get_n_solutions(cobj, ctypes.byref(n)))
        E = np.empty(max(1, n.value), dtype)
        _check(_lib.sqaod_rocm_dg_bf_searcher_get_E(cobj, _ptr(E), E.shape[0]))
        return E

    @staticmethod
    def prepare(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_bf_searcher_prepare(cobj))

    @staticmethod
    def calculate_E(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_bf_searcher_calculate_E(cobj))

    @staticmethod
    def make_solution(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_bf_searcher_make_solution(cobj))

    @staticmethod
    def search_range(cobj, dtype):
        done, cur = _i32(), _u64()
        _check(_lib.sqaod_rocm_dg_bf_searcher_search_range(cobj, ctypes.byref(done), ctypes.byref(cur)))
        return bool(done.value), cur.value

    @staticmethod
    def search(cobj, dtype):
        _check(_lib.sqaod_rocm_dg_bf_searcher_search(cobj))


class _BipartiteGraphBFSearcher:
    """rocm_bg_bf_searcher module: bf_searcher.inc:479-494 method table, bipartite form."""

    @staticmethod
    def new(dtype):
        out = _vp()
        _check(_lib.sqaod_rocm_bg_bf_searcher_new(_dt(dtype), ctypes.byref(out)))
        return out.value

    @staticmethod
    def delete(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_bf_searcher_delete(cobj))

    @staticmethod
    def assign_device(cobj, devobj, dtype):
        _check(_lib.sqaod_rocm_bg_bf_searcher_assign_device(cobj, devobj))

    @staticmethod
    def set_qubo(cobj, b0, b1, W, optimize, dtype):
        b0 = np.ascontiguousarray(b0, dtype)
        b1 = np.ascontiguousarray(b1, dtype)
        W = np.ascontiguousarray(W, dtype)
        if W.ndim != 2 or b0.ndim != 1 or b1.ndim != 1 or W.shape != (b1.shape[0], b0.shape[0]):
            raise RuntimeError("Shape mismatch: W must be N1 x N0 for b0[N0], b1[N1].")
        _check(_lib.sqaod_rocm_bg_bf_searcher_set_qubo(cobj, _ptr(b0), b0.shape[0], _ptr(b1), b1.shape[0], _ptr(W),
                                                      int(optimize)))

    @staticmethod
    def get_problem_size(cobj, dtype):
        n0, n1 = _i32(), _i32()
        _check(_lib.sqaod_rocm_bg_bf_searcher_get_problem_size(cobj, ctypes.byref(n0), ctypes.byref(n1)))
        return n0.value, n1.value

    @staticmethod
    def set_preferences(cobj, prefs, dtype):
        _set_prefs(_lib.sqaod_rocm_bg_bf_searcher_set_preference, cobj, prefs)

    @staticmethod
    def get_preferences(cobj, dtype):
        return _get_prefs(_lib.sqaod_rocm_bg_bf_searcher_get_preferences, cobj)

    @staticmethod
    def get_x(cobj, dtype):
        n = _i32()
        _check(_lib.sqaod_rocm_bg_bf_searcher_get_n_solutions(cobj, ctypes.byref(n)))
        N0, N1 = _BipartiteGraphBFSearcher.get_problem_size(cobj, dtype)
        x0 = np.empty((n.value, N0), np.int8)
        x1 = np.empty((n.value, N1), np.int8)
        _check(_lib.sqaod_rocm_bg_bf_searcher_get_x(cobj, _ptr(x0), _ptr(x1), n.value, N0, N1))
        return [(x0[k].copy(), x1[k].copy()) for k in range(n.value)]

    @staticmethod
    def get_E(cobj, dtype):
        n = _i32()
        _check(_lib.sqaod_rocm_bg_bf_searcher_get_n_solutions(cobj, ctypes.byref(n)))
        E = np.empty(max(1, n.value), dtype)
        _check(_lib.sqaod_rocm_bg_bf_searcher_get_E(cobj, _ptr(E), E.shape[0]))
        return E

    @staticmethod
    def prepare(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_bf_searcher_prepare(cobj))

    @staticmethod
    def calculate_E(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_bf_searcher_calculate_E(cobj))

    @staticmethod
    def make_solution(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_bf_searcher_make_solution(cobj))

    @staticmethod
    def search_range(cobj, dtype):
        done, c0, c1 = _i32(), _u64(), _u64()
        _check(_lib.sqaod_rocm_bg_bf_searcher_search_range(cobj, ctypes.byref(done), ctypes.byref(c0),
                                                          ctypes.byref(c1)))
        return bool(done.value), c0.value, c1.value

    @staticmethod
    def search(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_bf_searcher_search(cobj))


class _BipartiteGraphAnnealer:
    """rocm_bg_annealer module: annealer.inc:885-908 method table, bipartite form."""

    @staticmethod
    def new(dtype):
        out = _vp()
        _check(_lib.sqaod_rocm_bg_annealer_new(_dt(dtype), ctypes.byref(out)))
        return out.value

    @staticmethod
    def delete(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_delete(cobj))

    @staticmethod
    def assign_device(cobj, devobj, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_assign_device(cobj, devobj))

    @staticmethod
    def seed(cobj, seed, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_seed(cobj, int(seed) & 0xffffffffffffffff))

    @staticmethod
    def set_qubo(cobj, b0, b1, W, optimize, dtype):
        b0 = np.ascontiguousarray(b0, dtype).reshape(-1)
        b1 = np.ascontiguousarray(b1, dtype).reshape(-1)
        W = np.ascontiguousarray(W, dtype)
        if W.shape != (b1.shape[0], b0.shape[0]):
            raise RuntimeError("Shape mismatch: W must be N1 x N0 for b0[N0], b1[N1].")
        _check(_lib.sqaod_rocm_bg_annealer_set_qubo(cobj, _ptr(b0), b0.shape[0], _ptr(b1), b1.shape[0], _ptr(W),
                                                    int(optimize)))

    @staticmethod
    def set_hamiltonian(cobj, h0, h1, J, c, dtype):
        h0 = np.ascontiguousarray(h0, dtype).reshape(-1)
        h1 = np.ascontiguousarray(h1, dtype).reshape(-1)
        J = np.ascontiguousarray(J, dtype)
        cc = np.array([c], dtype)
        if J.shape != (h1.shape[0], h0.shape[0]):
            raise RuntimeError("Shape mismatch: J must be N1 x N0 for h0[N0], h1[N1].")
        _check(_lib.sqaod_rocm_bg_annealer_set_hamiltonian(cobj, _ptr(h0), h0.shape[0], _ptr(h1), h1.shape[0],
                                                           _ptr(J), _ptr(cc)))

    @staticmethod
    def get_problem_size(cobj, dtype):
        n0, n1 = _i32(), _i32()
        _check(_lib.sqaod_rocm_bg_annealer_get_problem_size(cobj, ctypes.byref(n0), ctypes.byref(n1)))
        return n0.value, n1.value

    @staticmethod
    def set_preferences(cobj, prefs, dtype):
        _set_prefs(_lib.sqaod_rocm_bg_annealer_set_preference, cobj, prefs)

    @staticmethod
    def get_preferences(cobj, dtype):
        return _get_prefs(_lib.sqaod_rocm_bg_annealer_get_preferences, cobj)

    @staticmethod
    def _m(cobj):
        m = _i32()
        _check(_lib.sqaod_rocm_bg_annealer_get_n_trotters(cobj, ctypes.byref(m)))
        return m.value

    @staticmethod
    def get_E(cobj, dtype):
        m = _BipartiteGraphAnnealer._m(cobj)
        E = np.empty(m, dtype)
        _check(_lib.sqaod_rocm_bg_annealer_get_E(cobj, _ptr(E), m))
        return E

    @staticmethod
    def get_E_fp64(cobj, dtype):
        m = _BipartiteGraphAnnealer._m(cobj)
        E = np.empty(m, np.float64)
        _check(_lib.sqaod_rocm_bg_annealer_get_E_fp64(cobj, _ptr(E), m))
        return E

    @staticmethod
    def _pairs(fn, cobj, dtype):
        m = _BipartiteGraphAnnealer._m(cobj)
        N0, N1 = _BipartiteGraphAnnealer.get_problem_size(cobj, dtype)
        a0 = np.empty((m, N0), np.int8)
        a1 = np.empty((m, N1), np.int8)
        _check(fn(cobj, _ptr(a0), _ptr(a1), m, N0, N1))
        return [(a0[i].copy(), a1[i].copy()) for i in range(m)]

    @staticmethod
    def get_x(cobj, dtype):
        return _BipartiteGraphAnnealer._pairs(_lib.sqaod_rocm_bg_annealer_get_x, cobj, dtype)

    @staticmethod
    def get_q(cobj, dtype):
        return _BipartiteGraphAnnealer._pairs(_lib.sqaod_rocm_bg_annealer_get_q, cobj, dtype)

    @staticmethod
    def get_hamiltonian(cobj, h0, h1, J, c, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_get_hamiltonian(cobj, _ptr(h0), _ptr(h1), _ptr(J), _ptr(c)))

    @staticmethod
    def set_q(cobj, qpair, dtype):
        q0 = np.ascontiguousarray(qpair[0], np.int8).reshape(-1)
        q1 = np.ascontiguousarray(qpair[1], np.int8).reshape(-1)
        _check(_lib.sqaod_rocm_bg_annealer_set_q(cobj, _ptr(q0), q0.shape[0], _ptr(q1), q1.shape[0]))

    @staticmethod
    def set_qset(cobj, qpairs, dtype):
        q0 = np.ascontiguousarray(np.stack([np.asarray(p[0], np.int8).reshape(-1) for p in qpairs]))
        q1 = np.ascontiguousarray(np.stack([np.asarray(p[1], np.int8).reshape(-1) for p in qpairs]))
        _check(_lib.sqaod_rocm_bg_annealer_set_qset(cobj, _ptr(q0), _ptr(q1), q0.shape[0], q0.shape[1], q1.shape[1]))

    @staticmethod
    def randomize_spin(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_randomize_spin(cobj))

    @staticmethod
    def calculate_E(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_calculate_E(cobj))

    @staticmethod
    def prepare(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_prepare(cobj))

    @staticmethod
    def make_solution(cobj, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_make_solution(cobj))

    @staticmethod
    def get_system_E(cobj, G, beta, dtype):
        E = _f64()
        _check(_lib.sqaod_rocm_bg_annealer_get_system_E(cobj, float(G), float(beta), ctypes.byref(E)))
        return np.dtype(dtype).type(E.value)

    @staticmethod
    def anneal_one_step(cobj, G, beta, dtype):
        _check(_lib.sqaod_rocm_bg_annealer_anneal_one_step(cobj, float(G), float(beta)))


class _Formulas:
    """rocm_formulas module: formulas.inc:586-606 method table.  Arrays are
    taken as given (the device-independent front end fixes types and shapes)."""

    @staticmethod
    def dg_formulas_new():
        out = _vp()
        _check(_lib.sqaod_rocm_formulas_new(ctypes.byref(out)))
        return out.value

    bg_formulas_new = dg_formulas_new

    @staticmethod
    def dg_formulas_delete(cobj):
        _check(_lib.sqaod_rocm_formulas_delete(cobj))

    bg_formulas_delete = dg_formulas_delete

    @staticmethod
    def dg_formulas_assign_device(cobj, devobj):
        _check(_lib.sqaod_rocm_formulas_assign_device(cobj, devobj))

    bg_formulas_assign_device = dg_formulas_assign_device

    @staticmethod
    def dense_graph_batch_calculate_E(cobj, E, W, x, dtype):
        _check(_lib.sqaod_rocm_dense_graph_batch_calculate_E(cobj, _dt(dtype), _ptr(E), _ptr(W), W.shape[0], _ptr(x),
                                                             x.shape[0]))

    @staticmethod
    def dense_graph_calculate_E(cobj, E, W, x, dtype):
        _Formulas.dense_graph_batch_calculate_E(cobj, E, W, x.reshape(1, -1), dtype)

    @staticmethod
    def dense_graph_calculate_hamiltonian(cobj, h, J, c, W, dtype):
        _check(_lib.sqaod_rocm_dense_graph_calculate_hamiltonian(cobj, _dt(dtype), _ptr(h), _ptr(J), _ptr(c), _ptr(W),
                                                                 W.shape[0]))

    @staticmethod
    def dense_graph_batch_calculate_E_from_spin(cobj, E, h, J, c, q, dtype):
        cc = np.array([c], dtype)
        _check(_lib.sqaod_rocm_dense_graph_batch_calculate_E_from_spin(cobj, _dt(dtype), _ptr(E), _ptr(h), _ptr(J),
                                                                       _ptr(cc), _ptr(q), J.shape[0], q.shape[0]))

    @staticmethod
    def dense_graph_calculate_E_from_spin(cobj, E, h, J, c, q, dtype):
        _Formulas.dense_graph_batch_calculate_E_from_spin(cobj, E, h, J, c, q.reshape(1, -1), dtype)

    @staticmethod
    def bipartite_graph_batch_calculate_E(cobj, E, b0, b1, W, x0, x1, dtype):
        _check(_lib.sqaod_rocm_bipartite_graph_batch_calculate_E(cobj, _dt(dtype), _ptr(E), _ptr(b0), b0.shape[0],
                                                                 _ptr(b1), b1.shape[0], _ptr(W), _ptr(x0), _ptr(x1),
                                                                 x0.shape[0]))

    @staticmethod
    def bipartite_graph_calculate_E(cobj, E, b0, b1, W, x0, x1, dtype):
        _Formulas.bipartite_graph_batch_calculate_E(cobj, E, b0, b1, W, x0.reshape(1, -1), x1.reshape(1, -1), dtype)

    @staticmethod
    def bipartite_graph_batch_calculate_E_2d(cobj, E, b0, b1, W, x0, x1, dtype):
        _check(_lib.sqaod_rocm_bipartite_graph_batch_calculate_E_2d(cobj, _dt(dtype), _ptr(E), _ptr(b0), b0.shape[0],
                                                                    _ptr(b1), b1.shape[0], _ptr(W), _ptr(x0),
                                                                    x0.shape[0], _ptr(x1), x1.shape[0]))

    @staticmethod
    def bipartite_graph_calculate_hamiltonian(cobj, h0, h1, J, c, b0, b1, W, dtype):
        _check(_lib.sqaod_rocm_bipartite_graph_calculate_hamiltonian(cobj, _dt(dtype), _ptr(h0), _ptr(h1), _ptr(J),
                                                                     _ptr(c), _ptr(b0), b0.shape[0], _ptr(b1),
                                                                     b1.shape[0], _ptr(W)))

    @staticmethod
    def bipartite_graph_batch_calculate_E_from_spin(cobj, E, h0, h1, J, c, q0, q1, dtype):
        cc = np.array([c], dtype)
        _check(_lib.sqaod_rocm_bipartite_graph_batch_calculate_E_from_spin(cobj, _dt(dtype), _ptr(E), _ptr(h0),
                                                                           h0.shape[0], _ptr(h1), h1.shape[0],
                                                                           _ptr(J), _ptr(cc), _ptr(q0), _ptr(q1),
                                                                           q0.shape[0]))

    @staticmethod
    def bipartite_graph_calculate_E_from_spin(cobj, E, h0, h1, J, c, q0, q1, dtype):
        _Formulas.bipartite_graph_batch_calculate_E_from_spin(cobj, E, h0, h1, J, c, q0.reshape(1, -1),
                                                              q1.reshape(1, -1), dtype)


class _DenseBatch:
    """ctypes front of sqaod_rocm_dense_batch_* (float32 instances)."""

    @staticmethod
    def new():
        p = _vp()
        _check(_lib.sqaod_rocm_dense_batch_new(ctypes.byref(p)))
        return p

    @staticmethod
    def delete(cobj):
        _check(_lib.sqaod_rocm_dense_batch_delete(cobj))

    @staticmethod
    def set_preferences(cobj, prefs):
        for k, v in prefs.items():
            _check(_lib.sqaod_rocm_dense_batch_set_preference(cobj, str(k).encode(), str(v).encode()))

    @staticmethod
    def add_qubo(cobj, W, m, seed, maximize):
        W = np.ascontiguousarray(W, np.float32)
        idx = _i32()
        _check(_lib.sqaod_rocm_dense_batch_add_qubo(cobj, _ptr(W), W.shape[0], int(m), int(seed) & ((1 << 64) - 1),
                                                    1 if maximize else 0, ctypes.byref(idx)))
        return idx.value

    @staticmethod
    def size(cobj):
        n = _i32()
        _check(_lib.sqaod_rocm_dense_batch_size(cobj, ctypes.byref(n)))
        return n.value

    @staticmethod
    def prepare(cobj):
        _check(_lib.sqaod_rocm_dense_batch_prepare(cobj))

    @staticmethod
    def anneal(cobj, G, beta, graph):
        G = np.ascontiguousarray(G, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        ms = _f64()
        _check(_lib.sqaod_rocm_dense_batch_anneal(cobj, _ptr(G), _ptr(beta), G.shape[0], 1 if graph else 0,
                                                  ctypes.byref(ms)))
        return ms.value

    @staticmethod
    def autotune(cobj, G, beta, steps):
        w, o = _i32(), _i32()
        _check(_lib.sqaod_rocm_dense_batch_autotune(cobj, float(G), float(beta), int(steps), ctypes.byref(w),
                                                    ctypes.byref(o)))
        return w.value, bool(o.value)

    @staticmethod
    def attempts_per_step(cobj):
        n = ctypes.c_longlong()
        _check(_lib.sqaod_rocm_dense_batch_attempts_per_step(cobj, ctypes.byref(n)))
        return n.value

    @staticmethod
    def shape(cobj, i):
        N, m = _i32(), _i32()
        _check(_lib.sqaod_rocm_dense_batch_get_shape(cobj, i, ctypes.byref(N), ctypes.byref(m)))
        return N.value, m.value

    @staticmethod
    def get_E(cobj, i):
        N, m = _DenseBatch.shape(cobj, i)
        E = np.empty(m, np.float32)
        _check(_lib.sqaod_rocm_dense_batch_get_E(cobj, i, _ptr(E), m))
        return E

    @staticmethod
    def get_x(cobj, i):
        N, m = _DenseBatch.shape(cobj, i)
        x = np.empty((m, N), np.int8)
        _check(_lib.sqaod_rocm_dense_batch_get_x(cobj, i, _ptr(x), m, N))
        return x

    @staticmethod
    def get_q(cobj, i):
        N, m = _DenseBatch.shape(cobj, i)
        q = np.empty((m, N), np.int8)
        _check(_lib.sqaod_rocm_dense_batch_get_q(cobj, i, _ptr(q), m, N))
        return q


class _Server:
    """ctypes front of sqaod_rocm_server_*."""

    @staticmethod
    def new():
        p = _vp()
        _check(_lib.sqaod_rocm_server_new(ctypes.byref(p)))
        return p

    @staticmethod
    def delete(cobj):
        _check(_lib.sqaod_rocm_server_delete(cobj))

    @staticmethod
    def set_preferences(cobj, prefs):
        for k, v in prefs.items():
            _check(_lib.sqaod_rocm_server_set_preference(cobj, str(k).encode(), str(v).encode()))

    @staticmethod
    def submit_dense(cobj, W, m, seed, G, beta):
        W = np.ascontiguousarray(W, np.float32)
        G = np.ascontiguousarray(G, np.float64)
        beta = np.ascontiguousarray(np.broadcast_to(beta, G.shape), np.float64)
        i = _i32()
        _check(_lib.sqaod_rocm_server_submit_dense(cobj, _ptr(W), W.shape[0], int(m), int(seed), _ptr(G), _ptr(beta),
                                                   G.shape[0], ctypes.byref(i)))
        return i.value

    @staticmethod
    def submit_bipartite(cobj, b0, b1, W, m, seed, G, beta):
        b0 = np.ascontiguousarray(b0, np.float32)
        b1 = np.ascontiguousarray(b1, np.float32)
        W = np.ascontiguousarray(W, np.float32)
        G = np.ascontiguousarray(G, np.float64)
        beta = np.ascontiguousarray(np.broadcast_to(beta, G.shape), np.float64)
        i = _i32()
        _check(_lib.sqaod_rocm_server_submit_bipartite(cobj, _ptr(b0), _ptr(b1), _ptr(W), b0.shape[0], b1.shape[0],
                                                       int(m), int(seed), _ptr(G), _ptr(beta), G.shape[0],
                                                       ctypes.byref(i)))
        return i.value

    @staticmethod
    def tick(cobj):
        r = _i32()
        _check(_lib.sqaod_rocm_server_tick(cobj, ctypes.byref(r)))
        return r.value

    @staticmethod
    def status(cobj):
        p, a, b, pk = _i32(), _i32(), _u64(), _u64()
        _check(_lib.sqaod_rocm_server_status(cobj, ctypes.byref(p), ctypes.byref(a), ctypes.byref(b),
                                             ctypes.byref(pk)))
        return {"pending": p.value, "active": a.value, "bytes_in_use": b.value, "peak_bytes": pk.value}

    @staticmethod
    def take_done(cobj, cap=4096):
        ids = np.empty(cap, np.int32)
        n = _i32()
        _check(_lib.sqaod_rocm_server_take_done(cobj, _ptr(ids), cap, ctypes.byref(n)))
        return [int(v) for v in ids[:n.value]]

    @staticmethod
    def result(cobj, i):
        kind, N0, N1, m = _i32(), _i32(), _i32(), _i32()
        _check(_lib.sqaod_rocm_server_result_shape(cobj, i, ctypes.byref(kind), ctypes.byref(N0), ctypes.byref(N1),
                                                   ctypes.byref(m)))
        E = np.empty(m.value, np.float32)
        x0 = np.empty((m.value, N0.value), np.int8)
        x1 = np.empty((m.value, N1.value), np.int8) if kind.value == 1 else None
        t = np.empty(3, np.float64)
        att = ctypes.c_longlong()
        _check(_lib.sqaod_rocm_server_result(cobj, i, _ptr(E), _ptr(x0), _ptr(x1) if x1 is not None else None,
                                             _ptr(t), ctypes.byref(att)))
        out = {"kind": "dense" if kind.value == 0 else "bipartite", "E": E, "submit_s": t[0], "admit_s": t[1],
               "finish_s": t[2], "attempts": att.value}
        if kind.value == 0:
            out["x"] = x0
        else:
            out["x0"], out["x1"] = x0, x1
        return out

    @staticmethod
    def error(cobj, i):
        """The failed job's message, or None for a job that completed."""
        buf = ctypes.create_string_buffer(1024)
        failed = _i32()
        _check(_lib.sqaod_rocm_server_result_error(cobj, i, buf, len(buf), ctypes.byref(failed)))
        return buf.value.decode() if failed.value else None

    @staticmethod
    def release(cobj, i):
        _check(_lib.sqaod_rocm_server_release(cobj, i))

    @staticmethod
    def now(cobj):
        t = _f64()
        _check(_lib.sqaod_rocm_server_now(cobj, ctypes.byref(t)))
        return t.value


rocm_device = _Device
rocm_dg_annealer = _DenseGraphAnnealer
rocm_dg_bf_searcher = _DenseGraphBFSearcher
rocm_bg_annealer = _BipartiteGraphAnnealer
rocm_bg_bf_searcher = _BipartiteGraphBFSearcher
rocm_formulas = _Formulas
rocm_dense_batch = _DenseBatch
rocm_server = _Server
