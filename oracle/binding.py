"""ctypes binding of librefcpu.so, the FP64 oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from molspin_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        L.ref_last_error.restype = C.c_char_p
        L.ref_dimension.restype = C.c_int64
        L.ref_trapezoid.restype = C.c_double
        L.ref_run_hsgeneral.argtypes = [P(abi.SystemDesc), P(abi.RunParams), C.c_int32, C.c_void_p,
                                        C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int]
        L.ref_run_columns.argtypes = [P(abi.SystemDesc), P(abi.RunParams), C.c_void_p, C.c_void_p,
                                      C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_void_p, C.c_int]
        L.ref_symmetric_decay.argtypes = [P(abi.SystemDesc), C.c_void_p, C.c_double, C.c_void_p]
        L.ref_time_infinity.argtypes = [P(abi.SystemDesc), P(abi.RunParams), C.c_void_p, C.c_void_p,
                                        C.c_int32, C.c_void_p]
        L.ref_lyapunov_integral.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_legacy_direct_yields.argtypes = [P(abi.SystemDesc), C.c_void_p, C.c_double, C.c_double,
                                               C.c_int, C.c_double, C.c_int, C.c_void_p]
        L.ref_legacy_stoch_yields.argtypes = [P(abi.SystemDesc), C.c_void_p, C.c_double, C.c_double,
                                              C.c_void_p, C.c_int32, C.c_int, C.c_void_p]
        L.ref_coherent_samples.argtypes = [P(abi.SystemDesc), C.c_int, C.c_uint64, C.c_int32, C.c_void_p]
        L.ref_suz_samples.argtypes = [P(abi.SystemDesc), C.c_int, C.c_uint64, C.c_int32, C.c_void_p]
        L.ref_fibonacci_point.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_void_p]
        L.ref_logspace.argtypes = [C.c_double, C.c_double, C.c_int32, C.c_void_p]
        L.ref_time_grid.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.ref_eig_hermitian.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_expm.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
        L.ref_trapezoid.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().ref_last_error().decode())


def _cplx(a, shape):
    return a.view(np.complex128).reshape(shape, order="F")


def dimension(packed):
    return int(lib().ref_dimension(packed.ref()))


def hamiltonian(packed, field=None):
    D = packed.dim
    out = np.zeros(2 * D * D)
    f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
    _check(lib().ref_hamiltonian_dense(packed.ref(), _p(f), _p(out)))
    return _cplx(out, (D, D))


def projector(packed, state):
    D = packed.dim
    out = np.zeros(2 * D * D)
    _check(lib().ref_state_projector_dense(packed.ref(), int(state), _p(out)))
    return _cplx(out, (D, D))


def reaction(packed):
    D = packed.dim
    out = np.zeros(2 * D * D)
    _check(lib().ref_reaction_dense(packed.ref(), _p(out)))
    return _cplx(out, (D, D))


def direct_factors(packed):
    D = packed.dim
    out = np.zeros(2 * D * D)
    n = C.c_int32()
    _check(lib().ref_direct_factors(packed.ref(), _p(out), C.byref(n)))
    return _cplx(out[: 2 * D * n.value], (D, n.value))


def time_grid(total, dt):
    n = C.c_int64()
    _check(lib().ref_time_grid(total, dt, C.byref(n), None))
    t = np.zeros(n.value)
    _check(lib().ref_time_grid(total, dt, C.byref(n), _p(t)))
    return t


def run_hsgeneral(packed, params, fields=None, columns=None, n_obs=None, threads=1, trace=False):
    """Returns (yields[n_items, n_obs], trace[n_items, n_samples, n_obs] or None)."""
    n_items = 1 if fields is None else len(fields)
    f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
    cols = None
    ncol = 0
    if columns is not None:
        cc = np.asfortranarray(columns, dtype=np.complex128)
        ncol = cc.shape[1]
        cols = cc.reshape(-1, order="F").view(np.float64).copy()
    if n_obs is None:
        n_obs = len(packed_transitions(packed)) if params.observables == abi.MSG_OBS_TRANSITION_YIELDS \
            else packed.desc.n_states
    ns = len(time_grid(params.total_time, params.time_step))
    y = np.zeros(n_items * n_obs)
    tr = np.zeros(n_items * ns * n_obs) if trace else None
    _check(lib().ref_run_hsgeneral(packed.ref(), C.byref(params), n_items, _p(f), _p(cols), ncol,
                                   _p(y), _p(tr), int(threads)))
    return y.reshape(n_items, n_obs), (tr.reshape(n_items, ns, n_obs) if trace else None)


def run_columns(packed, params, field=None, columns=None, col_range=None, max_steps=None,
                n_obs=None, threads=1):
    f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
    cols = None
    ncol = 0
    if columns is not None:
        cc = np.asfortranarray(columns, dtype=np.complex128)
        ncol = cc.shape[1]
        cols = cc.reshape(-1, order="F").view(np.float64).copy()
    if n_obs is None:
        n_obs = len(packed_transitions(packed)) if params.observables == abi.MSG_OBS_TRANSITION_YIELDS \
            else packed.desc.n_states
    ns = len(time_grid(params.total_time, params.time_step))
    if max_steps is None:
        max_steps = ns - 1
    nsm = min(ns, max_steps + 1)
    c0, c1 = col_range
    v = np.zeros(nsm * n_obs)
    _check(lib().ref_run_columns(packed.ref(), C.byref(params), _p(f), _p(cols), ncol, c0, c1,
                                 int(max_steps), _p(v), int(threads)))
    return v.reshape(nsm, n_obs)


def packed_transitions(packed):
    d = packed.desc
    return [i for i in range(d.n_transitions) if d.transitions[i].source_state >= 0]


def symmetric_decay(packed, k, field=None):
    f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
    y = np.zeros(packed.desc.n_states)
    _check(lib().ref_symmetric_decay(packed.ref(), _p(f), float(k), _p(y)))
    return y


def time_infinity(packed, params, field=None, columns=None):
    """HSGeneral method=timeinf yields of one item (the D^2 x D^2 time-integral solve)."""
    f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
    cols, ncol = None, 0
    if columns is not None:
        cc = np.asfortranarray(columns, dtype=np.complex128)
        ncol = cc.shape[1]
        cols = cc.reshape(-1, order="F").view(np.float64).copy()
    n_obs = len(packed_transitions(packed)) if params.observables == abi.MSG_OBS_TRANSITION_YIELDS \
        else packed.desc.n_states
    y = np.zeros(n_obs)
    _check(lib().ref_time_infinity(packed.ref(), C.byref(params), _p(f), _p(cols), ncol, _p(y)))
    return y


def lyapunov_integral(g, rho0):
    """X with G X + X G^dagger = -rho0 by the oracle's Kronecker-form solve."""
    n = g.shape[0]
    gg = np.asfortranarray(g, dtype=np.complex128).reshape(-1, order="F").view(np.float64).copy()
    rr = np.asfortranarray(rho0, dtype=np.complex128).reshape(-1, order="F").view(np.float64).copy()
    out = np.zeros(2 * n * n)
    _check(lib().ref_lyapunov_integral(n, _p(gg), _p(rr), _p(out)))
    return _cplx(out, (n, n))


def legacy_direct_yields(packed, total, dt, krylov=16, tol=0.0, correction=False, field=None):
    f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
    y = np.zeros(len(packed_transitions(packed)))
    _check(lib().ref_legacy_direct_yields(packed.ref(), _p(f), total, dt, krylov, tol, int(correction), _p(y)))
    return y


def legacy_stoch_yields(packed, total, dt, columns, correction=False, field=None):
    f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
    cc = np.asfortranarray(columns, dtype=np.complex128)
    cols = cc.reshape(-1, order="F").view(np.float64).copy()
    y = np.zeros(len(packed_transitions(packed)))
    _check(lib().ref_legacy_stoch_yields(packed.ref(), _p(f), total, dt, _p(cols), cc.shape[1],
                                         int(correction), _p(y)))
    return y


def coherent_samples(packed, state, seed, m):
    D = packed.dim
    out = np.zeros(2 * D * m)
    _check(lib().ref_coherent_samples(packed.ref(), int(state), int(seed), int(m), _p(out)))
    return _cplx(out, (D, m))


def suz_samples(packed, state, seed, m):
    D = packed.dim
    out = np.zeros(2 * D * m)
    _check(lib().ref_suz_samples(packed.ref(), int(state), int(seed), int(m), _p(out)))
    return _cplx(out, (D, m))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().ref_philox4x32_10(c, k, o)
    return list(o)


def fibonacci_point(n, i, magnitude=1.0):
    o = np.zeros(3)
    lib().ref_fibonacci_point(n, i, magnitude, _p(o))
    return o


def logspace(a, b, n):
    o = np.zeros(n)
    lib().ref_logspace(a, b, n, _p(o))
    return o


def trapezoid(times, values):
    t = np.ascontiguousarray(times, dtype=np.float64)
    v = np.ascontiguousarray(values, dtype=np.float64)
    return lib().ref_trapezoid(_p(t), _p(v), len(t))


def eig_hermitian(a):
    n = a.shape[0]
    aa = np.asfortranarray(a, dtype=np.complex128).reshape(-1, order="F").view(np.float64).copy()
    w = np.zeros(n)
    v = np.zeros(2 * n * n)
    _check(lib().ref_eig_hermitian(n, _p(aa), _p(w), _p(v)))
    return w, _cplx(v, (n, n))


def expm(a):
    n = a.shape[0]
    aa = np.asfortranarray(a, dtype=np.complex128).reshape(-1, order="F").view(np.float64).copy()
    o = np.zeros(2 * n * n)
    _check(lib().ref_expm(n, _p(aa), _p(o)))
    return _cplx(o, (n, n))
