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


def test_struct_layouts_match_header_sizes():
    # sizes of the POD structs as the C compiler lays them out (natural alignment)
    assert C.sizeof(abi.SpinDesc) == 16 + 72
    assert C.sizeof(abi.InteractionDesc) == 40 + 8 + 24 + 72
    assert C.sizeof(abi.RunParams) == 24 + 16 + 8 + 8


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
