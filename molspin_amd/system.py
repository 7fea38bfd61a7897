"""Structural spin-system description in MolSpin's object model.

`SpinSystem` collects Spin / Interaction / State / Transition / relaxation
objects with MolSpin's meaning (SpinAPI/Spin.cpp, SpinAPI/Interaction.cpp,
SpinAPI/State.cpp, SpinAPI/Transition.cpp) and packs them into the POD
`msg_system_desc` of include/molspin_gpu.h.  The packed form keeps its ctypes
buffers alive for as long as the `Packed` object lives.
"""
import ctypes as C
import math

from . import abi


def _tensor(spec):
    """(isotropic flag, 3x3 lab-frame matrix) from a number, a 3-vector of
    principal values on the laboratory axes (Tensor 'anisotropic(a b c)' plus an
    isotropic part) or a 3x3 matrix (Tensor.cpp:216-300, LabFrame :426-433)."""
    if isinstance(spec, (int, float)):
        m = [[float(spec) if i == j else 0.0 for j in range(3)] for i in range(3)]
        return True, m
    if isinstance(spec, dict):
        iso = float(spec.get("isotropic", 0.0))
        ani = spec.get("anisotropic", (0.0, 0.0, 0.0))
        mat = spec.get("matrix")
        m = [[iso if i == j else 0.0 for j in range(3)] for i in range(3)]
        if mat is not None:
            for i in range(3):
                for j in range(3):
                    m[i][j] += float(mat[i][j])
        for i in range(3):
            m[i][i] += float(ani[i])
    else:
        rows = list(spec)
        if len(rows) == 3 and not isinstance(rows[0], (list, tuple)):
            m = [[float(rows[i]) if i == j else 0.0 for j in range(3)] for i in range(3)]
        else:
            m = [[float(rows[i][j]) for j in range(3)] for i in range(3)]
    tr = (m[0][0] + m[1][1] + m[2][2]) / 3.0
    dev = sum(abs(m[i][j] - (tr if i == j else 0.0)) for i in range(3) for j in range(3))
    return dev < 1e-10, m  # IsIsotropic (Tensor.cpp:834-840)


class SpinSystem:
    def __init__(self, name="RP"):
        self.name = name
        self.spins = []          # dicts: name, mult, electron, tensor
        self.interactions = []
        self.states = []         # dicts: name, spins(list of idx), amps(list complex)
        self.transitions = []    # dicts: name, source(idx), rate
        self.initial = []        # state indices
        self.weights = None
        self.relaxations = []    # dicts: spin1, spin2, rate

    # -- objects -----------------------------------------------------------
    def spin_index(self, name):
        for i, s in enumerate(self.spins):
            if s["name"] == name:
                return i
        raise KeyError(name)

    def add_spin(self, name, spin="1/2", electron=False, tensor=None):
        """Spin object; spin as "1/2", "1", 0.5 ...; default tensor isotropic 2.0
        (Spin.cpp:29)."""
        if isinstance(spin, str):
            if "/" in spin:
                a, b = spin.split("/")
                s = float(a) / float(b)
            else:
                s = float(spin)
        else:
            s = float(spin)
        mult = int(round(2 * s + 1))
        iso, m = _tensor(2.0 if tensor is None else tensor)
        self.spins.append(dict(name=name, mult=mult, electron=bool(electron), tensor=m, iso=iso))
        return len(self.spins) - 1

    def _idx(self, x):
        return x if isinstance(x, int) else self.spin_index(x)

    def add_zeeman(self, spins, field, prefactor=1.0, common_prefactor=True,
                   ignore_tensors=False, swept=False, name=None, time_dependence=None,
                   frequency=1.0, phase=0.0, axis=(0.0, 0.0, 1.0), perpendicular=False):
        """Zeeman interaction; time_dependence None, "linearpolarized" or
        "circularpolarized" (fieldtype, Interaction.cpp:500-560) with frequency (rad/ns),
        phase (rad), axis and perpendicularoscillations."""
        td = {None: abi.MSG_TD_NONE, "linearpolarized": abi.MSG_TD_LINEAR,
              "circularpolarized": abi.MSG_TD_CIRCULAR}[time_dependence]
        self.interactions.append(dict(
            type=abi.MSG_INTERACTION_ZEEMAN, g1=[self._idx(s) for s in spins], g2=[],
            field=[float(x) for x in field], tensor=None, prefactor=float(prefactor),
            common=common_prefactor, ignore=ignore_tensors, swept=swept, name=name,
            td=td, freq=float(frequency), phase=float(phase), axis=[float(x) for x in axis],
            perp=bool(perpendicular)))

    def add_hyperfine(self, group1, group2, tensor=None, prefactor=1.0, common_prefactor=True,
                      ignore_tensors=False, name=None, exchange=False):
        t = None if tensor is None else _tensor(tensor)
        self.interactions.append(dict(
            type=abi.MSG_INTERACTION_EXCHANGE if exchange else abi.MSG_INTERACTION_DOUBLESPIN,
            g1=[self._idx(s) for s in group1], g2=[self._idx(s) for s in group2],
            field=[0.0, 0.0, 0.0], tensor=t, prefactor=float(prefactor), common=common_prefactor,
            ignore=ignore_tensors, swept=False, name=name))

    def add_exchange(self, group1, group2, J, prefactor=1.0, common_prefactor=True, name=None):
        self.add_hyperfine(group1, group2, tensor=J, prefactor=prefactor,
                           common_prefactor=common_prefactor, name=name, exchange=True)

    def add_state(self, name, spins, amps):
        """State over `spins` (listed order, first most significant) with
        amplitudes `amps` (len = product of multiplicities)."""
        idx = [self._idx(s) for s in spins]
        n = 1
        for i in idx:
            n *= self.spins[i]["mult"]
        amps = [complex(a) for a in amps]
        if len(amps) != n:
            raise ValueError("state %s: %d amplitudes for dimension %d" % (name, len(amps), n))
        self.states.append(dict(name=name, spins=idx, amps=amps))
        return len(self.states) - 1

    def state_index(self, name):
        for i, s in enumerate(self.states):
            if s["name"] == name:
                return i
        raise KeyError(name)

    def add_transition(self, name, source, rate):
        src = source if isinstance(source, int) else self.state_index(source)
        self.transitions.append(dict(name=name, source=src, rate=float(rate)))

    def set_initial(self, states, weights=None):
        self.initial = [s if isinstance(s, int) else self.state_index(s) for s in states]
        self.weights = None if weights is None else [float(w) for w in weights]

    def add_st_dephasing(self, spin1, spin2, rate):
        self.relaxations.append(dict(spin1=self._idx(spin1), spin2=self._idx(spin2), rate=float(rate)))

    def add_radical_pair_states(self, e1, e2):
        """S, T0, T+, T- over the two electrons (Example/magnetoreception/...)."""
        r = 1.0 / math.sqrt(2.0)
        self.add_state("S", [e1, e2], [0, r, -r, 0])
        self.add_state("T0", [e1, e2], [0, r, r, 0])
        self.add_state("Tp", [e1, e2], [1, 0, 0, 0])
        self.add_state("Tm", [e1, e2], [0, 0, 0, 1])

    def dimension(self):
        d = 1
        for s in self.spins:
            d *= s["mult"]
        return d

    # -- packing -----------------------------------------------------------
    def pack(self):
        return Packed(self)


class Packed:
    """Owns the ctypes arrays behind one msg_system_desc."""

    def __init__(self, sysobj):
        s = sysobj
        pool = []
        ns = len(s.spins)
        self.spins = (abi.SpinDesc * max(ns, 1))()
        for i, sp in enumerate(s.spins):
            d = self.spins[i]
            d.multiplicity = sp["mult"]
            d.is_electron = int(sp["electron"])
            d.tensor_isotropic = int(sp["iso"])
            for a in range(3):
                for b in range(3):
                    d.tensor[3 * a + b] = sp["tensor"][a][b]
        ni = len(s.interactions)
        self.inter = (abi.InteractionDesc * max(ni, 1))()
        for i, it in enumerate(s.interactions):
            d = self.inter[i]
            d.type = it["type"]
            d.group1_offset = len(pool); d.group1_count = len(it["g1"]); pool += it["g1"]
            d.group2_offset = len(pool); d.group2_count = len(it["g2"]); pool += it["g2"]
            d.common_prefactor = int(bool(it["common"]))
            d.ignore_tensors = int(bool(it["ignore"]))
            d.field_swept = int(bool(it["swept"]))
            d.prefactor = it["prefactor"]
            for a in range(3):
                d.field[a] = it["field"][a]
            d.time_dependence = it.get("td", abi.MSG_TD_NONE)
            d.td_frequency = it.get("freq", 1.0)
            d.td_phase = it.get("phase", 0.0)
            d.td_perpendicular = int(it.get("perp", False))
            for a, x in enumerate(it.get("axis", (0.0, 0.0, 1.0))):
                d.td_axis[a] = x
            if it["tensor"] is not None:
                iso, m = it["tensor"]
                d.tensor_present = 1
                d.tensor_isotropic = int(iso)
                for a in range(3):
                    for b in range(3):
                        d.tensor[3 * a + b] = m[a][b]
        nst = len(s.states)
        self.states = (abi.StateDesc * max(nst, 1))()
        amps = []
        for i, st in enumerate(s.states):
            d = self.states[i]
            d.spins_offset = len(pool); d.spins_count = len(st["spins"]); pool += st["spins"]
            d.amp_offset = len(amps)
            amps += st["amps"]
        nt = len(s.transitions)
        self.trans = (abi.TransitionDesc * max(nt, 1))()
        for i, t in enumerate(s.transitions):
            self.trans[i].source_state = t["source"]
            self.trans[i].rate = t["rate"]
        self.initial = (C.c_int32 * max(len(s.initial), 1))(*s.initial)
        self.weights = None if s.weights is None else (C.c_double * len(s.weights))(*s.weights)
        nr = len(s.relaxations)
        self.relax = (abi.RelaxationDesc * max(nr, 1))()
        for i, r in enumerate(s.relaxations):
            self.relax[i].spin1 = r["spin1"]; self.relax[i].spin2 = r["spin2"]; self.relax[i].rate = r["rate"]
        self.pool = (C.c_int32 * max(len(pool), 1))(*pool)
        flat = []
        for a in amps:
            flat += [a.real, a.imag]
        self.amps = (C.c_double * max(len(flat), 1))(*flat)
        self.desc = abi.SystemDesc(
            ns, self.spins, ni, self.inter, nst, self.states, nt, self.trans,
            len(s.initial), self.initial,
            C.cast(self.weights, C.POINTER(C.c_double)) if self.weights is not None else None,
            nr, self.relax, len(pool), self.pool, len(amps), self.amps)
        self.dim = s.dimension()

    def ref(self):
        return C.byref(self.desc)
