"""CPU-side checks of the C ABI: the library loads, exports every entry point the
public header declares, and host-only entry points behave without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from molspin_amd import abi, backend

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:msg_status|const char \*)\s*(msg_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = header_entry_points()
    assert len(names) >= 15
    L = backend.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(backend.EXPORTS)


STRUCTS = [("msg_spin_desc", "SpinDesc"), ("msg_interaction_desc", "InteractionDesc"),
           ("msg_state_desc", "StateDesc"), ("msg_transition_desc", "TransitionDesc"),
           ("msg_relaxation_desc", "RelaxationDesc"), ("msg_system_desc", "SystemDesc"),
           ("msg_run_params", "RunParams"), ("msg_plan", "Plan"), ("msg_launch_shape", "LaunchShape"),
           ("msg_uneven_entry", "UnevenEntry")]


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives the
    header's struct (gcc on include/molspin_gpu.h)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "molspin_gpu.h"', 'int main(void) {']
    for cname, pyname in STRUCTS:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (pyname, cname))
        for f in getattr(abi, pyname)._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (pyname, f[0], cname, f[0]))
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in out:
        if not line:
            continue
        pyname, field, value = line.split()
        cls = getattr(abi, pyname)
        got = C.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert got == int(value), (pyname, field, got, value)


def test_time_grid_without_device():
    t = backend.time_grid(0.25, 0.2)
    assert list(t) == [0.0, 0.2, 0.25]
    t = backend.time_grid(5000.0, 1.0)
    assert len(t) == 5001 and t[-1] == 5000.0


def test_context_creation_fails_cleanly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(backend.BackendError) as e:
        backend.Context(0)
    assert e.value.status in (abi.MSG_DEVICE_ERROR, abi.MSG_INVALID_ARG)
