"""The C-ABI library loads and exports every symbol include/nqes.h declares (CPU only)."""
import ctypes
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "nqes.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nqes_[A-Za-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    import nqes
    lib = nqes.load_library()
    names = declared_functions()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/nqes.h but not exported"
    assert sorted(nqes.EXPORTS) == names


def test_version_and_sizing_need_no_device():
    import nqes
    lib = nqes.load_library()
    assert lib.nqes_version().startswith(b"nqes-mi355x")
    # O of the paper-size problem in FP64 complex: S * P * 16 bytes dominates
    b = lib.nqes_required_bytes(32, 64, 2, 4096, 256, 0)
    assert b >= (1 << 20) * 4288 * 16
    assert lib.nqes_required_bytes(0, 64, 2, 4096, 256, 0) == -1
    # mixed, factored S*v: spins and c*tanh(theta) rows (2^21 rows x 144 complex64),
    # far below the 2^20 x 4288 FP32 O of the streamed form
    mb = lib.nqes_required_bytes(32, 64, 2, 4096, 256, 1)
    assert (1 << 21) * 144 * 8 <= mb < (1 << 20) * 4288 * 4
