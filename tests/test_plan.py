"""Plan resolution (msg_resolve_plan, host only) against ResolveExecutionPlan's
behaviour: defaults (HSExecutionPlan.h:33-117), key aliases, and the rejection
messages exercised by the reference's own tests (tests_HSGeneral.cpp:760-820,
1000-1001, 1370-1390, 1773-1808).  Plans the reference accepts but that select
physics off the device path resolve to MSG_UNSUPPORTED."""
import pytest

from molspin_amd import abi, backend


def resolve(text):
    return backend.resolve_plan(text)


def rejects(text, fragment, status=abi.MSG_INVALID_ARG):
    with pytest.raises(backend.BackendError) as e:
        resolve("type=HSGeneral;" + text)
    assert e.value.status == status, str(e.value)
    assert fragment in str(e.value), str(e.value)


def test_defaults_are_the_reference_defaults():
    """tests_HSGeneral.cpp:761-773: static, timeevolution, direct, full, exponential."""
    p = resolve("type=hsgeneral;")
    r = p.params
    assert r.calculation == abi.MSG_CALC_TIMEEVOLUTION
    assert r.sampling == abi.MSG_SAMPLING_DIRECT
    assert r.method == abi.MSG_METHOD_EXPONENTIAL and p.propagation == abi.MSG_PROP_EXPONENTIAL
    assert r.observables == abi.MSG_OBS_STATE_POPULATIONS
    assert (r.total_time, r.time_step, r.krylov_size, r.yield_corrections, r.seed) == (10000.0, 1.0, 16, 0, 1)
    assert (p.montecarlo_samples, p.sampling_method, p.auto_seed, p.device_backend, p.time_infinity) == \
        (1, abi.MSG_TRACE_SUZ, 1, 0, 0)


def test_config_task_blocks():
    p = resolve("type=HSGeneral; calculation=yields; sampling=direct; propagationmethod=krylov; krylovsize=16;"
                " totaltime=5000; timestep=1; transitionyields=true; yieldcorrections=false; backend=mi355x;")
    r = p.params
    assert (r.calculation, r.observables, r.method, r.krylov_size, r.total_time, r.time_step) == \
        (abi.MSG_CALC_YIELDS, abi.MSG_OBS_TRANSITION_YIELDS, abi.MSG_METHOD_KRYLOV, 16, 5000.0, 1.0)
    assert p.device_backend == 1
    p = resolve("type=HSGeneral; dynamics=static; calculation=timeevolution; sampling=stochastic;"
                " propagationmethod=autoexpm; precision=double; samplingmethod=suz; montecarlosamples=8192;"
                " autoseed=false; seed=431; timestep=0.025; totaltime=2;")  # st_dephasing.msd, stochastic task
    r = p.params
    assert r.sampling == abi.MSG_SAMPLING_GIVEN and r.method == abi.MSG_METHOD_EXPONENTIAL
    assert p.propagation == abi.MSG_PROP_AUTOEXPM and p.montecarlo_samples == 8192 and p.auto_seed == 0
    assert r.seed == 431 and r.time_step == 0.025 and r.total_time == 2.0
    p = resolve("type=HSGeneral; calculation=yield; tracesampling=mc; samplingmethod=coherent; krylovsize=0;")
    assert p.params.calculation == abi.MSG_CALC_YIELDS and p.params.sampling == abi.MSG_SAMPLING_GIVEN
    assert p.sampling_method == abi.MSG_TRACE_COHERENT and p.params.krylov_size == 16
    p = resolve("type=HSGeneral; calculation=yields; method=timeinf;")
    assert p.time_infinity == 1


def test_legacy_task_types():
    p = resolve("type=statichs-direct-yields; totaltime=300; timestep=0.5; yieldcorrections=true;")
    assert p.params.task == abi.MSG_TASK_LEGACY_DIRECT_YIELDS and p.params.yield_corrections == 1
    p = resolve("type=StaticHS-Stoch-Yields; montecarlosamples=64; samplingmethod=coherent;")
    assert p.params.task == abi.MSG_TASK_LEGACY_STOCH_YIELDS and p.montecarlo_samples == 64
    assert p.params.sampling == abi.MSG_SAMPLING_GIVEN
    assert resolve("type=statichs-symmetricdecay;").params.task == abi.MSG_TASK_SYMMETRIC_DECAY
    assert resolve("type=StaticHS;").params.task == abi.MSG_TASK_SYMMETRIC_DECAY
    # krylovtol of the legacy direct-yields task (TaskStaticHSDirectYields.cpp:373-410)
    assert resolve("type=statichs-direct-yields; krylovtol=1e-9;").params.krylov_tol == 1e-9
    assert resolve("type=statichs-direct-yields;").params.krylov_tol == 0.0
    with pytest.raises(backend.BackendError) as e:
        resolve("type=StaticSS;")
    assert e.value.status == abi.MSG_UNSUPPORTED


def test_reference_rejections():
    rejects("calculation=spectra;", "standalone task")
    rejects("calculation=resonance;", "standalone task")
    rejects("calculation=timeevolution;approximation=rwa;", "not a static Hamiltonian secular approximation")
    rejects("calculation=timeevolution;powderaveraging=true;", "requires a generated powder grid")
    rejects("dynamics=dynamic;calculation=yields;method=timeinf;", "dynamic yields require finite-time")
    rejects("calculation=yields;relaxationmodel=nakajimazwanzig;", "Nakajima-Zwanzig")
    rejects("calculation=yields;approximation=secular;", "requires hamiltonianh0list")
    rejects("calculation=yields;reactionoperators=lindblad;", "superspace reaction-superoperator")
    rejects("calculation=yields;reactionoperators=unknown;", "must be haberkorn")
    rejects("calculation=yields;powdergammapoints=4;", "requires a generated theta/phi powder grid")
    rejects("calculation=timeevolution;sampling=stochastic;samplingmethod=unknown;", "samplingmethod must be suz or coherent")
    rejects("calculation=timeevolution;observables=states;spinlist=E;", "cannot be combined with spinlist")
    rejects("calculation=timeevolution;observables=spins;", "requires spinlist")
    rejects("calculation=timeevolution;observables=transitionyields;", "requires calculation=yields")
    rejects("calculation=yields;observables=cidsp;", "requires spinlist")
    rejects("calculation=timeevolution;observables=unknown;", "observables must be")
    rejects("dynamics=static;calculation=timeevolution;sampling=direct;powdersamplingpoints=3;", "requires hamiltonianh0list")
    rejects("calculation=timeevolution;hamiltonianh0list=H0;powdergrid=unknown;powdersamplingpoints=3;",
            "powdergrid must be uniform, sophe, or octant")
    rejects("timestep=0;", "timestep must be finite and positive")
    rejects("totaltime=-1;", "totaltime must be finite and non-negative")
    rejects("propagationmethod=foo;", "propagationmethod must be normal/exp, autoexpm, krylov, or rk4")
    rejects("montecarlosamples=0;", "montecarlosamples must be at least one")
    rejects("krylovtol=-1;", "krylovtol must be finite and positive")
    rejects("calculation=timeevolution;method=timeinf;", "requires method=timeevo")
    rejects("backend=cuda;", "backend must be cpu or mi355x")


def test_valid_but_off_the_device_path():
    p = resolve("type=HSGeneral;propagationmethod=uncoupled;calculation=yields;")
    assert p.params.method == abi.MSG_METHOD_UNCOUPLED
    p = resolve("type=HSGeneral;propagationmethod=rk4;")
    assert p.params.method == abi.MSG_METHOD_RK4 and p.propagation == abi.MSG_PROP_RK4
    p = resolve("type=HSGeneral;dynamics=timedependent;calculation=yields;")
    assert p.params.dynamics == 1
    assert resolve("type=HSGeneral;").params.dynamics == 0
    for text in ("calculation=timeevolution;observables=spins;spinlist=E;",
                 "calculation=yields;approximation=secular;hamiltonianh0list=H0;",
                 "calculation=timeevolution;hamiltonianh0list=H0;powdergrid=uniform;powdersamplingpoints=8;",
                 "krylovsize=64;", 'pulsesequence=["p1 10"];'):
        rejects(text, "not on the device path", abi.MSG_UNSUPPORTED)


def test_property_splitting_follows_object_parser():
    """';' inside parentheses stays in the value; keys are case-insensitive;
    empty keys or values are dropped; unparsable numbers read as absent."""
    p = resolve("TYPE=HSGeneral;; timestep=abc; totaltime=12; weird(1;2)=x; calculation=Yields")
    assert p.params.calculation == abi.MSG_CALC_YIELDS
    assert p.params.time_step == 1.0 and p.params.total_time == 12.0
