"""MSD front-end (msg_msd_compile): MSD text -> descriptor, per-step swept fields
and task plans.

CPU: the compiled system equals the same system built through molspin_amd.system
(oracle Hamiltonian, projectors, reaction operator and direct factors); the step
fields follow the RunSection step loop (FibonacciSphere vectors pinned to
tests_actions.cpp through workloads.fibonacci_point, RotateVector closed form,
LogSpace on a prefactor, first/period scheduling); objects off the path are refused.
GPU: a compiled Fibonacci sweep runs as one batch and matches the builder's batch
and the oracle.
"""
import math

import numpy as np
import pytest

from molspin_amd import abi, backend, workloads
from oracle import binding as ref


def rp_msd(nuc, hf, steps=1, actions="", field="0 0 0.05", tasks=None):
    """MSD text of a radical pair: electrons E1/E2 (g 2.0023), nuclei `nuc` as
    (name, spin, electron), hyperfine specs `hf` (MSD tensor strings, mT), Zeeman
    `zee` on both electrons, S/T0/T+/T- with sinks at 1e-3 /ns, initial singlet."""
    lines = ["// radical pair written for the front-end tests", "SpinSystem rp", "{",
             "  Spin E1 { type = electron; spin = 1/2; tensor = isotropic(2.0023); }",
             "  Spin E2 { type = electron; spin = 1/2; tensor = isotropic(2.0023); }"]
    for name, spin, _ in nuc:
        lines.append("  Spin %s { type = nucleus; spin = %s; tensor = isotropic(1); }" % (name, spin))
    for (name, _, el), t in zip(nuc, hf):
        lines.append("  Interaction hf_%s { type = hyperfine; group1 = %s; group2 = %s; tensor = %s; prefactor = 0.001; }"
                     % (name, el, name, t))
    lines += ['  Interaction zee { type = zeeman; spins = E1, E2; field = "%s"; prefactor = 0.001; }' % field,
              "  /* the four electron-pair states */",
              "  State S { spins(E1,E2) = |1/2,-1/2> - |-1/2,1/2>; }",
              "  State T0 { spins(E1,E2) = |1/2,-1/2> + |-1/2,1/2>; }",
              "  State Tp { spin(E1) = |1/2>; spin(E2) = |1/2>; }",
              "  State Tm { spins(E1,E2) = |-1/2,-1/2>; }"]
    for st in ("S", "T0", "Tp", "Tm"):
        lines.append("  Transition k%s { type = sink; sourcestate = %s; rate = 0.001; }" % (st, st))
    lines += ["  Properties properties { initialstate = S; }", "}",
              "Settings", "{", "  Settings general { steps = %d; }" % steps, actions, "}", "Run", "{"]
    for name, body in (tasks or [("yields", "type = HSGeneral; calculation = yields; propagationmethod = krylov; "
                                             "totaltime = 3000; timestep = 2;")]):
        lines.append("  Task %s { %s }" % (name, body))
    lines.append("}")
    return "\n".join(lines)


def config1_msd(**kw):
    nuc = [("A0", "1/2", "E1"), ("A1", "1/2", "E1"), ("B0", "1", "E2")]
    return rp_msd(nuc, ["isotropic(0.5)", "isotropic(1.2)", "isotropic(0.8)"], **kw)


def matrix_spec(m):
    return 'matrix("%s")' % ";".join(" ".join(repr(float(x)) for x in row) for row in m)


def config2_msd(**kw):
    """config 2's nuclei and anisotropic hyperfine tensors (same seeded stream)."""
    rng = workloads.SplitMix64(2, 0)
    hf = [workloads._aniso_hf(rng) for _ in range(10)]
    nuc = [("A%d" % i, "1/2", "E1") for i in range(6)] + [("B%d" % i, "1/2", "E2") for i in range(4)]
    return rp_msd(nuc, [matrix_spec(m) for m in hf], **kw)


def same_system(a, b, fa=None, fb=None, full=True):
    """Oracle matrices of two descriptors agree (the Hamiltonian only unless `full`)."""
    assert a.dim == b.dim
    assert np.abs(ref.hamiltonian(a, fa) - ref.hamiltonian(b, fb)).max() < 1e-13
    if not full:
        return
    for st in range(4):
        assert np.abs(ref.projector(a, st) - ref.projector(b, st)).max() < 1e-14
    assert np.abs(ref.reaction(a) - ref.reaction(b)).max() < 1e-16
    Ba, Bb = ref.direct_factors(a), ref.direct_factors(b)
    assert np.abs(Ba @ Ba.conj().T - Bb @ Bb.conj().T).max() < 1e-15


def test_config1_from_msd_equals_the_builder():
    m = backend.Msd(config1_msd())
    s, _ = workloads.config1()
    same_system(m, s.pack())
    assert m.steps == 1 and m.fields is None
    assert len(m.tasks) == 1 and m.tasks[0]["status"] == abi.MSG_OK and m.tasks[0]["name"] == "yields"
    prm = m.tasks[0]["plan"].params
    assert (prm.calculation, prm.method, prm.total_time, prm.time_step) == (abi.MSG_CALC_YIELDS, abi.MSG_METHOD_KRYLOV,
                                                                          3000.0, 2.0)


def test_fibonacci_sweep_fields():
    """FibonacciSphere on zee.field with points = steps: step i uses point i (point 0
    installed before step 1), prefactor folded in; the compiled Hamiltonian at each
    step's field equals config 2's at its Fibonacci field."""
    n = 16
    act = "  Action sphere { type = FibonacciSphere; vector = rp.zee.field; points = %d; }" % n
    m = backend.Msd(config2_msd(steps=n, actions=act))
    assert m.steps == n and m.fields.shape == (n, 3)
    s, cfg = workloads.config2(n_orientations=n)
    want = 0.001 * np.array(cfg["fields"])
    # the front-end takes cos/sin of the float angle in float (std::cos(float), as
    # RetrievePoint does); numpy's float32 cos may differ by one ulp
    assert np.abs(m.fields - want).max() <= 2e-7 * 0.001 * 0.05
    p = s.pack()
    for i in (0, 5, n - 1):
        same_system(m, p, m.fields[i], m.fields[i] / 0.001, full=(i == 0))


def test_rotate_vector_and_schedule():
    """RotateVector by 1 degree about y turns z towards x; with first = 3, period = 2
    the rotation happens before steps 3, 5, 7, ... only."""
    act = "  Action turn { type = RotateVector; vector = rp.zee.field; axis = \"0 1 0\"; value = 1; }"
    m = backend.Msd(config1_msd(steps=10, actions=act))
    ang = np.radians(np.arange(10))
    want = 0.001 * 0.05 * np.stack([np.sin(ang), 0 * ang, np.cos(ang)], axis=1)
    assert np.abs(m.fields - want).max() < 1e-17
    act = ("  Action turn { type = RotateVector; vector = rp.zee.field; axis = \"0 1 0\"; value = 10; "
           "first = 3; period = 2; }")
    m = backend.Msd(config1_msd(steps=8, actions=act))
    turns = np.array([0, 0, 1, 1, 2, 2, 3, 3])
    ang = np.radians(10 * turns)
    want = 0.001 * 0.05 * np.stack([np.sin(ang), 0 * ang, np.cos(ang)], axis=1)
    assert np.abs(m.fields - want).max() < 1e-17


def test_logspace_prefactor_sweep():
    """LogSpace on zee.prefactor: config 4's 64 field strengths (1 uT .. 10 mT along z)."""
    act = "  Action strength { type = LogSpace; scalar = rp.zee.prefactor; points = 64; min = -3; max = 1; }"
    m = backend.Msd(config1_msd(steps=64, actions=act, field="0 0 0.001"))
    want = np.array([(0.0, 0.0, 0.001 * b) for b in workloads.logspace(-3.0, 1.0, 64)])
    assert np.abs(m.fields - want).max() <= 1e-15 * np.abs(want).max()


def test_tensor_and_ket_forms():
    """isotropic + anisotropic along given axes, matrix(), and ket coefficients
    (i, -i, numbers) over spin-1 and spin-1/2 bases."""
    text = """SpinSystem t {
      Spin E1 { type = electron; tensor = 2.0 + anisotropic(0.1, 0.2, -0.3) + axis1(0, 1, 0) + axis2(-1, 0, 0) + axis3(0, 0, 1); }
      Spin E2 { type = electron; tensor = matrix("2.1 0.1 0;0.1 2.0 0;0 0 1.9"); }
      Spin N { type = nucleus; spin = 1; tensor = isotropic(1); }
      State A { spins(E1,N) = 0.5|1/2,1> + i|-1/2,0> - 2i|-1/2,-1>; }
      State B { spin(N) = -i|0>; spins(E1,E2) = |1/2,-1/2>; }
      Properties p { initialstate = A; }
    }"""
    m = backend.Msd(text)
    d = m.desc
    t1 = np.array(d.spins[0].tensor[:]).reshape(3, 3)
    assert np.allclose(t1, np.diag([2.2, 2.1, 1.7]), atol=1e-15) and not d.spins[0].tensor_isotropic
    t2 = np.array(d.spins[1].tensor[:]).reshape(3, 3)
    assert np.allclose(t2, [[2.1, 0.1, 0], [0.1, 2.0, 0], [0, 0, 1.9]], atol=1e-15)
    assert d.spins[2].multiplicity == 3 and d.spins[0].multiplicity == 2
    amps = np.array(d.amp_pool[:2 * d.amp_pool_len]).view(np.complex128)
    a = amps[d.states[0].amp_offset:d.states[0].amp_offset + 6]
    assert np.allclose(a, [0.5, 0, 0, 0, 1j, -2j])
    b = amps[d.states[1].amp_offset:d.states[1].amp_offset + 12]
    want = np.kron([0, -1j, 0], [0, 1, 0, 0])
    assert np.allclose(b, want)
    assert [d.index_pool[d.states[1].spins_offset + k] for k in range(3)] == [2, 0, 1]


@pytest.mark.parametrize("text,status,words", [
    ("SpinSystem a { Spin e { type = electron; } Pulse p { type = instantpulse; } }", abi.MSG_UNSUPPORTED, "pulse"),
    ("SpinSystem a { Spin e { type = electron; } Interaction d { type = dipole; group1 = e; group2 = e; } }",
     abi.MSG_UNSUPPORTED, "dipole"),
    ("SpinSystem a { Spin e { type = electron; } Operator o { type = relaxationt1; spins = e; rate = 1; } }",
     abi.MSG_UNSUPPORTED, "relaxationdephasing"),
    ("SpinSystem a { Spin e { type = electron; } }  SpinSystem b { Spin e { type = electron; } }",
     abi.MSG_UNSUPPORTED, "one SpinSystem"),
    ("SpinSystem a { Spin e { type = electron; } Interaction z { type = zeeman; spins = e; field = \"0 0 1\"; } }"
     "Settings { Action x { type = ScaleVector; vector = a.z.field; value = 2; } }", abi.MSG_UNSUPPORTED, "ScaleVector"),
    ("SpinSystem a { Spin e { type = electron; } State s { spins(e) = |1/2>; } Properties p { initialstate = q; } }",
     abi.MSG_INVALID_ARG, "unknown state"),
    ("SpinSystem a { Spin e { type = electron; spin = 1/2 }", abi.MSG_INVALID_ARG, "unterminated"),
    ("#include \"other.msd\"", abi.MSG_UNSUPPORTED, "directives"),
])
def test_refusals(text, status, words):
    with pytest.raises(backend.BackendError) as e:
        backend.Msd(text)
    assert e.value.status == status and words.lower() in str(e.value).lower()


def test_tasks_report_their_plans():
    """Every Run task goes through the plan resolver; tasks off the path keep the
    resolver's status and message."""
    tasks = [("a", "type = HSGeneral; calculation = yields; method = timeinf;"),
             ("b", "type = StaticHS-SymmetricDecay;"),
             ("c", "type = HSGeneral; calculation = timeevolution; powdersamplingpoints = 3; hamiltonianh0list = H0;")]
    m = backend.Msd(config1_msd(tasks=tasks))
    st = [t["status"] for t in m.tasks]
    assert [t["name"] for t in m.tasks] == ["a", "b", "c"]
    assert st[0] == abi.MSG_OK and m.tasks[0]["plan"].time_infinity == 1
    assert st[1] == abi.MSG_OK and m.tasks[1]["plan"].params.task == abi.MSG_TASK_SYMMETRIC_DECAY
    assert st[2] == abi.MSG_UNSUPPORTED


@pytest.mark.gpu
def test_device_runs_a_compiled_sweep():
    """A compiled 8-point Fibonacci sweep of config 2 runs as one msg_propagate batch
    (short grid) and gives the builder's batch and the oracle's yields."""
    n = 8
    act = "  Action sphere { type = FibonacciSphere; vector = rp.zee.field; points = %d; }" % n
    m = backend.Msd(config2_msd(steps=n, actions=act, tasks=[
        ("y", "type = HSGeneral; calculation = yields; propagationmethod = krylov; totaltime = 40; timestep = 2;")]))
    ctx = backend.Context(0)
    try:
        prm = m.tasks[0]["plan"].params
        y = ctx.upload(m).propagate(prm, fields=m.fields)[0]
        s, cfg = workloads.config2(n_orientations=n)
        yb = ctx.upload(s.pack()).propagate(prm, fields=cfg["fields"])[0]
        print("compiled", y[:, 0], "builder", yb[:, 0])
        assert np.abs(y - yb).max() < 1e-9
        yo, _ = ref.run_hsgeneral(m, prm, fields=m.fields[:2], threads=8)
        assert np.abs(y[:2] - yo).max() <= 1e-5
    finally:
        ctx.close()
