"""Python binding of libmolspin_gpu.so (the C ABI in include/molspin_gpu.h).

The library is built in-tree by molspin_amd.build; this module refuses to run
without it (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmolspin_gpu.so")
_LIB = None

EXPORTS = [
    "msg_ctx_create", "msg_ctx_destroy", "msg_last_error", "msg_system_upload", "msg_system_release",
    "msg_system_dimension", "msg_query_footprint", "msg_direct_columns", "msg_observable_count",
    "msg_time_grid", "msg_propagate", "msg_step", "msg_server_create", "msg_server_destroy",
    "msg_server_submit", "msg_server_pump", "msg_server_collect", "msg_last_kernel_time",
    "msg_trace_samples", "msg_symmetric_decay", "msg_resolve_plan", "msg_autotune", "msg_time_infinity",
    "msg_msd_compile", "msg_msd_system", "msg_msd_steps", "msg_msd_task_count", "msg_msd_task", "msg_msd_destroy",
    "msg_propagate_uneven", "msg_propagate_noalloc", "msg_server_stats",
]


class BackendError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, str(status)), message))
        self.status = status


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmolspin_gpu.so is not built (python -m molspin_amd.build)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        vp = C.c_void_p
        L.msg_last_error.restype = C.c_char_p
        L.msg_last_error.argtypes = [vp]
        L.msg_ctx_create.argtypes = [C.c_int, P(vp)]
        L.msg_ctx_destroy.argtypes = [vp]
        L.msg_system_upload.argtypes = [vp, P(abi.SystemDesc), P(vp)]
        L.msg_system_release.argtypes = [vp, vp]
        L.msg_system_dimension.argtypes = [vp, P(C.c_int64)]
        L.msg_query_footprint.argtypes = [vp, P(abi.RunParams), C.c_int32, C.c_int32, P(C.c_int64)]
        L.msg_direct_columns.argtypes = [vp, P(C.c_int32)]
        L.msg_observable_count.argtypes = [vp, P(abi.RunParams), P(C.c_int32)]
        L.msg_time_grid.argtypes = [C.c_double, C.c_double, P(C.c_int64), vp]
        L.msg_propagate.argtypes = [vp, vp, P(abi.RunParams), C.c_int32, vp, vp, C.c_int32, vp, vp]
        L.msg_step.argtypes = [vp, vp, P(abi.RunParams), vp, C.c_double, C.c_int64, vp, C.c_int32]
        L.msg_server_create.argtypes = [vp, C.c_int32, C.c_int64, P(vp)]
        L.msg_server_destroy.argtypes = [vp]
        L.msg_server_submit.argtypes = [vp, vp, P(abi.RunParams), C.c_int32, vp, P(C.c_int64)]
        L.msg_server_pump.argtypes = [vp, C.c_int32, P(C.c_int32)]
        L.msg_server_collect.argtypes = [vp, C.c_int64, vp, P(C.c_int32)]
        L.msg_server_stats.argtypes = [vp, P(C.c_int64), P(C.c_int64)]
        L.msg_last_kernel_time.argtypes = [vp, P(C.c_double), P(C.c_int32)]
        L.msg_trace_samples.argtypes = [vp, C.c_int32, C.c_int32, C.c_uint64, C.c_int32, vp]
        L.msg_symmetric_decay.argtypes = [vp, vp, C.c_int32, vp, C.c_double, vp]
        L.msg_time_infinity.argtypes = [vp, vp, P(abi.RunParams), C.c_int32, vp, vp, C.c_int32, vp]
        L.msg_msd_compile.argtypes = [C.c_char_p, P(vp), C.c_char_p, C.c_int32]
        L.msg_msd_system.argtypes = [vp, P(P(abi.SystemDesc))]
        L.msg_msd_steps.argtypes = [vp, P(C.c_int32), P(C.c_int32), vp]
        L.msg_msd_task_count.argtypes = [vp, P(C.c_int32)]
        L.msg_msd_task.argtypes = [vp, C.c_int32, P(abi.Plan), C.c_char_p, C.c_int32, C.c_char_p, C.c_int32]
        L.msg_msd_destroy.argtypes = [vp]
        L.msg_resolve_plan.argtypes = [C.c_char_p, P(abi.Plan), C.c_char_p, C.c_int32]
        L.msg_autotune.argtypes = [vp, vp, P(abi.RunParams), C.c_int32, vp, vp, C.c_int32, C.c_int32,
                                   P(abi.LaunchShape)]
        L.msg_propagate_uneven.argtypes = [vp, C.c_int32, P(abi.UnevenEntry)]
        L.msg_propagate_noalloc.argtypes = [vp, vp, P(abi.RunParams), C.c_int32, vp, vp, C.c_int32, vp, vp]
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int if name != "msg_last_error" else C.c_char_p
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _columns(columns):
    """Column-major complex128 D x M as a flat float64 array (no copy when the
    input already has that layout)."""
    cc = np.asfortranarray(columns, dtype=np.complex128)
    return cc.reshape(-1, order="F").view(np.float64), cc.shape[1]


class Msd:
    """A compiled MSD file (msg_msd_compile): `packed`-like system handle for
    Context.upload, the per-step swept fields and the tasks with their plans."""

    def __init__(self, text):
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        st = lib().msg_msd_compile(text.encode(), C.byref(h), err, 512)
        if st != abi.MSG_OK:
            raise BackendError(st, err.value.decode())
        self.h = h
        d = C.POINTER(abi.SystemDesc)()
        lib().msg_msd_system(self.h, C.byref(d))
        self.desc = d.contents
        n, sw = C.c_int32(), C.c_int32()
        lib().msg_msd_steps(self.h, C.byref(n), C.byref(sw), None)
        self.steps = n.value
        self.fields = None
        if sw.value:
            f = np.zeros(3 * n.value)
            lib().msg_msd_steps(self.h, C.byref(n), C.byref(sw), _ptr(f))
            self.fields = f.reshape(-1, 3)
        cnt = C.c_int32()
        lib().msg_msd_task_count(self.h, C.byref(cnt))
        self.tasks = []
        for i in range(cnt.value):
            plan = abi.Plan()
            name = C.create_string_buffer(256)
            why = C.create_string_buffer(512)
            st = lib().msg_msd_task(self.h, i, C.byref(plan), name, 256, why, 512)
            self.tasks.append(dict(name=name.value.decode(), status=st, plan=plan, why=why.value.decode()))
        d = self.desc
        self.dim = 1
        for i in range(d.n_spins):
            self.dim *= d.spins[i].multiplicity

    def ref(self):
        return C.byref(self.desc)

    def close(self):
        if self.h:
            lib().msg_msd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def resolve_plan(task_properties):
    """msg_resolve_plan: an MSD Task property list -> abi.Plan (host only).
    Raises BackendError with the resolver's message."""
    plan = abi.Plan()
    err = C.create_string_buffer(512)
    st = lib().msg_resolve_plan(task_properties.encode(), C.byref(plan), err, len(err))
    if st != abi.MSG_OK:
        raise BackendError(st, err.value.decode())
    return plan


class Context:
    """One HIP device context with its own stream (msg_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().msg_ctx_create(int(device), C.byref(h))
        if st != abi.MSG_OK:
            raise BackendError(st, lib().msg_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            lib().msg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st):
        if st != abi.MSG_OK:
            raise BackendError(st, lib().msg_last_error(self.h).decode())

    def upload(self, packed):
        return System(self, packed)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def propagate_uneven(self, entries):
        """msg_propagate_uneven: several systems of different dimension in one
        persistent launch.  `entries`: dicts with keys system, params and optionally
        fields, columns, trace (bool); returns a list of (yields[n_items, n_obs],
        trace[n_items, n_samples, n_obs] or None)."""
        arr = (abi.UnevenEntry * len(entries))()
        keep, outs = [], []
        for k, ent in enumerate(entries):
            system, params = ent["system"], ent["params"]
            fields, columns = ent.get("fields"), ent.get("columns")
            n_items = 1 if fields is None else len(fields)
            f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
            cols, ncol = (None, 0) if columns is None else _columns(columns)
            nobs = system.n_obs(params)
            ns = len(time_grid(params.total_time, params.time_step))
            y = np.zeros(n_items * nobs)
            tr = np.zeros(n_items * ns * nobs) if ent.get("trace") else None
            keep += [params, f, cols, y, tr]
            arr[k] = abi.UnevenEntry(system.h, C.pointer(params), n_items, _ptr(f), _ptr(cols), ncol, _ptr(y), _ptr(tr))
            outs.append((y.reshape(n_items, nobs), None if tr is None else tr.reshape(n_items, ns, nobs)))
        self.check(lib().msg_propagate_uneven(self.h, len(entries), arr))
        return outs

    def last_kernel_ms(self):
        """Device time of the propagation kernel of the last propagate call."""
        ms = C.c_double()
        n = C.c_int32()
        self.check(lib().msg_last_kernel_time(self.h, C.byref(ms), C.byref(n)))
        return ms.value


class System:
    """An uploaded spin system (msg_system)."""

    def __init__(self, ctx, packed):
        self.ctx = ctx
        self.packed = packed
        h = C.c_void_p()
        ctx.check(lib().msg_system_upload(ctx.h, packed.ref(), C.byref(h)))
        self.h = h
        d = C.c_int64()
        ctx.check(lib().msg_system_dimension(h, C.byref(d)))
        self.dim = d.value

    def release(self):
        if self.h:
            lib().msg_system_release(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def n_obs(self, params):
        n = C.c_int32()
        self.ctx.check(lib().msg_observable_count(self.h, C.byref(params), C.byref(n)))
        return n.value

    def trace_samples(self, state, m, seed, method=0):
        """Unit-norm trace samples of State `state` (D x m, MolSpin order); method 0
        coherent, 1 SU(Z)."""
        out = np.zeros(2 * self.dim * m)
        self.ctx.check(lib().msg_trace_samples(self.h, int(state), int(method), int(seed), int(m), _ptr(out)))
        return out.view(np.complex128).reshape((self.dim, m), order="F")

    def direct_columns(self):
        n = C.c_int32()
        self.ctx.check(lib().msg_direct_columns(self.h, C.byref(n)))
        return n.value

    def footprint(self, params, n_items, n_columns=0):
        b = C.c_int64()
        self.ctx.check(lib().msg_query_footprint(self.h, C.byref(params), int(n_items), int(n_columns), C.byref(b)))
        return b.value

    def propagate(self, params, fields=None, columns=None, trace=False):
        """Returns (yields[n_items, n_obs], trace[n_items, n_samples, n_obs] or None)."""
        n_items = 1 if fields is None else len(fields)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
        cols, ncol = (None, 0) if columns is None else _columns(columns)
        nobs = self.n_obs(params)
        ns = len(time_grid(params.total_time, params.time_step))
        y = np.zeros(n_items * nobs)
        tr = np.zeros(n_items * ns * nobs) if trace else None
        self.ctx.check(lib().msg_propagate(self.ctx.h, self.h, C.byref(params), n_items, _ptr(f), _ptr(cols), ncol,
                                           _ptr(y), _ptr(tr)))
        return y.reshape(n_items, nobs), (tr.reshape(n_items, ns, nobs) if trace else None)

    def propagate_noalloc(self, params, fields=None, columns=None, trace=False):
        """msg_propagate_noalloc: as propagate, with every device buffer in the
        library's fixed scratch (no device allocation)."""
        n_items = 1 if fields is None else len(fields)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
        cols, ncol = (None, 0) if columns is None else _columns(columns)
        nobs = self.n_obs(params)
        ns = len(time_grid(params.total_time, params.time_step))
        y = np.zeros(n_items * nobs)
        tr = np.zeros(n_items * ns * nobs) if trace else None
        self.ctx.check(lib().msg_propagate_noalloc(self.ctx.h, self.h, C.byref(params), n_items, _ptr(f), _ptr(cols),
                                                   ncol, _ptr(y), _ptr(tr)))
        return y.reshape(n_items, nobs), (tr.reshape(n_items, ns, nobs) if trace else None)

    def autotune(self, params, fields=None, columns=None, sample_steps=4):
        """msg_autotune: times the candidate launch shapes on this sample and keeps
        the fastest for the context; returns the chosen abi.LaunchShape."""
        n_items = 1 if fields is None else len(fields)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
        cols, ncol = (None, 0) if columns is None else _columns(columns)
        shape = abi.LaunchShape()
        self.ctx.check(lib().msg_autotune(self.ctx.h, self.h, C.byref(params), n_items, _ptr(f), _ptr(cols), ncol,
                                          int(sample_steps), C.byref(shape)))
        return shape

    def symmetric_decay(self, fields=None, rate=0.0):
        """Infinite-time yields of every State under symmetric decay
        (TaskStaticHSSymmetricDecay); returns [n_items, n_states]."""
        n_items = 1 if fields is None else len(fields)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
        y = np.zeros(n_items * self.packed.desc.n_states)
        self.ctx.check(lib().msg_symmetric_decay(self.ctx.h, self.h, n_items, _ptr(f), float(rate), _ptr(y)))
        return y.reshape(n_items, -1)

    def time_infinity(self, params, fields=None, columns=None):
        """HSGeneral method=timeinf yields (msg_time_infinity); returns [n_items, n_obs]."""
        n_items = 1 if fields is None else len(fields)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
        cols, ncol = (None, 0) if columns is None else _columns(columns)
        y = np.zeros(n_items * self.n_obs(params))
        self.ctx.check(lib().msg_time_infinity(self.ctx.h, self.h, C.byref(params), n_items, _ptr(f), _ptr(cols),
                                               ncol, _ptr(y)))
        return y.reshape(n_items, -1)

    def step(self, params, factors, dt, step_index=0, field=None):
        """HSPropagator::StepStochastic equivalent; returns the new factors."""
        cols, ncol = _columns(factors)
        cols = cols.copy()  # msg_step writes the factors in place; keep the caller's array
        f = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
        self.ctx.check(lib().msg_step(self.ctx.h, self.h, C.byref(params), _ptr(f), float(dt), int(step_index),
                                      _ptr(cols), ncol))
        return cols.view(np.complex128).reshape((self.dim, ncol), order="F")


class Server:
    """Continuous-batching yield server (msg_server)."""

    def __init__(self, ctx, max_slots, hbm_budget_bytes):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().msg_server_create(ctx.h, int(max_slots), int(hbm_budget_bytes), C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            lib().msg_server_destroy(self.h)
            self.h = None

    def submit(self, system, params, fields):
        f = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1)
        rid = C.c_int64()
        self.ctx.check(lib().msg_server_submit(self.h, system.h, C.byref(params), len(f) // 3, _ptr(f), C.byref(rid)))
        self._keep[rid.value] = (system, params, len(f) // 3, system.n_obs(params))
        return rid.value

    def pump(self, max_steps):
        n = C.c_int32()
        self.ctx.check(lib().msg_server_pump(self.h, int(max_steps), C.byref(n)))
        return n.value

    def stats(self):
        """(shared launches, requests they held) since the server was created."""
        a, b = C.c_int64(), C.c_int64()
        self.ctx.check(lib().msg_server_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def collect(self, rid):
        system, params, nf, nobs = self._keep[rid]
        y = np.zeros(nf * nobs)
        done = C.c_int32()
        self.ctx.check(lib().msg_server_collect(self.h, int(rid), _ptr(y), C.byref(done)))
        if done.value:
            del self._keep[rid]
            return y.reshape(nf, nobs)
        return None


def time_grid(total, dt):
    n = C.c_int64()
    if lib().msg_time_grid(float(total), float(dt), C.byref(n), None) != abi.MSG_OK:
        raise BackendError(abi.MSG_INVALID_ARG, "invalid time grid")
    t = np.zeros(n.value)
    lib().msg_time_grid(float(total), float(dt), C.byref(n), _ptr(t))
    return t
