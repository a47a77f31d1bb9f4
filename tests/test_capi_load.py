"""The shared library loads on a CPU-only host and exports every function of include/dftbhip.h."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    import ctypes
    from dftbplus_amd import api
    assert os.path.exists(api.LIBPATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(api.LIBPATH)
    header = open(os.path.join(ROOT, "include", "dftbhip.h")).read()
    names = set(re.findall(r"^int (dftbhip_\w+)\(", header, flags=re.M))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(api.EXPORTS) <= names


def test_no_device_is_reported_cleanly():
    from dftbplus_amd import api
    import ctypes
    n = ctypes.c_int(-1)
    assert api.lib().dftbhip_get_gpus_available(ctypes.byref(n)) == 0
    assert n.value >= 0
