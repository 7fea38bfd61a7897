"""ctypes mirror of include/molspin_gpu.h (the C ABI of the device backend).

Only plain data types live here; loading the HIP library is done by
molspin_amd.backend, which fails loudly when the library is missing.
"""
import ctypes as C

MSG_OK = 0
MSG_INVALID_ARG = 1
MSG_UNSUPPORTED = 2
MSG_OOM_ADMISSION = 3
MSG_DEVICE_ERROR = 4
MSG_NUMERIC_BREAKDOWN = 5

STATUS_NAMES = {
    0: "MSG_OK", 1: "MSG_INVALID_ARG", 2: "MSG_UNSUPPORTED", 3: "MSG_OOM_ADMISSION",
    4: "MSG_DEVICE_ERROR", 5: "MSG_NUMERIC_BREAKDOWN",
}

MSG_INTERACTION_ZEEMAN = 1
MSG_INTERACTION_DOUBLESPIN = 2
MSG_INTERACTION_EXCHANGE = 3

MSG_CALC_YIELDS = 0
MSG_CALC_TIMEEVOLUTION = 1
MSG_OBS_TRANSITION_YIELDS = 0
MSG_OBS_STATE_POPULATIONS = 1
MSG_METHOD_AUTO = 0
MSG_METHOD_KRYLOV = 1
MSG_METHOD_EXPONENTIAL = 2
MSG_METHOD_RK4 = 3
MSG_METHOD_UNCOUPLED = 4
MSG_SAMPLING_DIRECT = 0
MSG_SAMPLING_GIVEN = 1
MSG_TASK_HSGENERAL = 0
MSG_TASK_LEGACY_DIRECT_YIELDS = 1
MSG_TASK_LEGACY_STOCH_YIELDS = 2
MSG_TASK_SYMMETRIC_DECAY = 3
MSG_PROP_EXPONENTIAL = 0
MSG_PROP_AUTOEXPM = 1
MSG_PROP_KRYLOV = 2
MSG_PROP_RK4 = 3
MSG_TD_NONE = 0
MSG_TD_LINEAR = 1
MSG_TD_CIRCULAR = 2
MSG_TRACE_COHERENT = 0
MSG_TRACE_SUZ = 1


class SpinDesc(C.Structure):
    _fields_ = [("multiplicity", C.c_int32), ("is_electron", C.c_int32),
                ("tensor_isotropic", C.c_int32), ("reserved", C.c_int32),
                ("tensor", C.c_double * 9)]


class InteractionDesc(C.Structure):
    _fields_ = [("type", C.c_int32),
                ("group1_offset", C.c_int32), ("group1_count", C.c_int32),
                ("group2_offset", C.c_int32), ("group2_count", C.c_int32),
                ("tensor_present", C.c_int32), ("tensor_isotropic", C.c_int32),
                ("common_prefactor", C.c_int32), ("ignore_tensors", C.c_int32),
                ("field_swept", C.c_int32),
                ("prefactor", C.c_double), ("field", C.c_double * 3), ("tensor", C.c_double * 9),
                ("time_dependence", C.c_int32), ("td_perpendicular", C.c_int32),
                ("td_frequency", C.c_double), ("td_phase", C.c_double), ("td_axis", C.c_double * 3)]


class StateDesc(C.Structure):
    _fields_ = [("spins_offset", C.c_int32), ("spins_count", C.c_int32),
                ("amp_offset", C.c_int32), ("reserved", C.c_int32)]


class TransitionDesc(C.Structure):
    _fields_ = [("source_state", C.c_int32), ("reserved", C.c_int32), ("rate", C.c_double)]


class RelaxationDesc(C.Structure):
    _fields_ = [("spin1", C.c_int32), ("spin2", C.c_int32), ("rate", C.c_double)]


class SystemDesc(C.Structure):
    _fields_ = [("n_spins", C.c_int32), ("spins", C.POINTER(SpinDesc)),
                ("n_interactions", C.c_int32), ("interactions", C.POINTER(InteractionDesc)),
                ("n_states", C.c_int32), ("states", C.POINTER(StateDesc)),
                ("n_transitions", C.c_int32), ("transitions", C.POINTER(TransitionDesc)),
                ("n_initial", C.c_int32), ("initial_states", C.POINTER(C.c_int32)),
                ("initial_weights", C.POINTER(C.c_double)),
                ("n_relaxations", C.c_int32), ("relaxations", C.POINTER(RelaxationDesc)),
                ("index_pool_len", C.c_int32), ("index_pool", C.POINTER(C.c_int32)),
                ("amp_pool_len", C.c_int32), ("amp_pool", C.POINTER(C.c_double))]


class RunParams(C.Structure):
    _fields_ = [("calculation", C.c_int32), ("observables", C.c_int32),
                ("method", C.c_int32), ("krylov_size", C.c_int32),
                ("sampling", C.c_int32), ("yield_corrections", C.c_int32),
                ("total_time", C.c_double), ("time_step", C.c_double),
                ("seed", C.c_uint64), ("task", C.c_int32), ("dynamics", C.c_int32),
                ("krylov_tol", C.c_double)]


class Plan(C.Structure):
    _fields_ = [("params", RunParams), ("montecarlo_samples", C.c_int32), ("sampling_method", C.c_int32),
                ("auto_seed", C.c_int32), ("device_backend", C.c_int32), ("time_infinity", C.c_int32),
                ("propagation", C.c_int32)]


class LaunchShape(C.Structure):
    _fields_ = [("path", C.c_int32), ("groups_per_thread", C.c_int32), ("threads", C.c_int32),
                ("slots", C.c_int32), ("kernel_ms", C.c_double)]


class UnevenEntry(C.Structure):
    _fields_ = [("system", C.c_void_p), ("params", C.POINTER(RunParams)), ("n_items", C.c_int32),
                ("fields", C.c_void_p), ("columns", C.c_void_p), ("n_columns", C.c_int32),
                ("yields", C.c_void_p), ("trace", C.c_void_p)]


def run_params(calculation=MSG_CALC_YIELDS, observables=MSG_OBS_TRANSITION_YIELDS,
               method=MSG_METHOD_KRYLOV, krylov_size=16, sampling=MSG_SAMPLING_DIRECT,
               yield_corrections=False, total_time=10000.0, time_step=1.0, seed=1, task=MSG_TASK_HSGENERAL,
               dynamics=False, krylov_tol=0.0):
    """RunParams with HSExecutionPlan defaults (HSExecutionPlan.h:60-79)."""
    return RunParams(calculation, observables, method, krylov_size, sampling,
                     int(bool(yield_corrections)), float(total_time), float(time_step), int(seed), int(task),
                     int(bool(dynamics)), float(krylov_tol))
