"""The C-ABI library loads on a CPU-only host and exports every symbol that
include/dndsr_amd.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    hdr = open(os.path.join(ROOT, "include", "dndsr_amd.h")).read()
    return sorted(set(re.findall(r"\b(dndsr_[a-z0-9_]+)\s*\(", hdr)))


def test_header_declares_the_boundary():
    names = _declared()
    for n in ("dndsr_create", "dndsr_vr_construct", "dndsr_vr_iter", "dndsr_eval_rhs", "dndsr_eval_dt",
              "dndsr_lusgs_init", "dndsr_update_sgs", "dndsr_update_sgs_f32", "dndsr_ssprk3"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from dndsr_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole header too
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_error_path_without_context():
    from dndsr_amd import _lib
    L = _lib.lib()
    assert L.dndsr_sync(None) != 0
    assert b"null context" in L.dndsr_last_error()


def test_kernel_id_count_matches_header():
    """dndsr_get_timing writes DNDSR_NUM_KERNEL_IDS entries: the Python list sizes the buffers."""
    from dndsr_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "dndsr_amd.h")).read()
    n = int(re.search(r"#define\s+DNDSR_NUM_KERNEL_IDS\s+(\d+)", hdr).group(1))
    assert len(_lib.KERNEL_IDS) == n


def test_cpp_host_mirror_driver_is_built():
    """The C++ host-mirror driver links against the library and runs its argument check
    without touching the GPU."""
    import subprocess
    from dndsr_amd import build as b
    assert os.path.exists(b.MIRROR_TEST_BIN)
    r = subprocess.run([b.MIRROR_TEST_BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
