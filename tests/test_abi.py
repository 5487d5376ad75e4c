"""The C-ABI library loads without a GPU and exports every entry point
include/foort_gpu.h declares; the ctypes mirror matches the header layout."""
import ctypes as C
import os
import re

import foort_amd
from foort_amd import abi as A

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "foort_gpu.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(foort_\w+)\s*\(", src)))


def test_exports_every_declared_symbol():
    L = foort_amd.lib()
    names = declared_functions()
    assert "foort_gpu_open" in names and "foort_gpu_trace_device" in names
    for n in names:
        assert hasattr(L, n), n
    assert sorted(foort_amd.EXPORTS) == names


def test_descriptor_layout_and_defaults():
    L = foort_amd.lib()
    assert L.foort_desc_size() == C.sizeof(A.RunDesc)
    d = A.RunDesc()
    L.foort_desc_init.argtypes = [C.POINTER(A.RunDesc)]
    L.foort_desc_init(C.byref(d))
    ref = A.default_desc()
    for name, _ in A.RunDesc._fields_:
        a, b = getattr(d, name), getattr(ref, name)
        if isinstance(a, (A.UpdateFrequency,)):
            assert (a.update_n_steps, a.update_start, a.update_finish) == (b.update_n_steps, b.update_start, b.update_finish)
        elif hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def test_open_errors_are_reported_without_gpu():
    L = foort_amd.lib()
    d = A.default_desc()
    d.integrator = 9
    h = C.c_void_p()
    assert L.foort_gpu_open(C.byref(d), C.byref(h)) == -1
    assert b"integrator" in L.foort_gpu_last_error(None)
