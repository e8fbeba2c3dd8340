"""The C-ABI libraries load, export every symbol include/les_gpu.h declares,
and the ctypes struct mirrors match the compiled struct sizes (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "les_gpu.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(les_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["les_gpu_create", "les_gpu_step", "les_gpu_rhs", "les_gpu_upload_state", "les_gpu_download_state"]:
        assert n in names


def test_engine_library_exports_all(lesgpu):
    L = C.CDLL(lesgpu.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes_match(lesgpu, abi):
    out = (C.c_size_t * 6)()
    L = lesgpu.lib()
    L.les_abi_struct_sizes.argtypes = [C.POINTER(C.c_size_t)]
    assert L.les_abi_struct_sizes(out) == 0
    assert out[0] == C.sizeof(abi.Params)
    assert out[1] == C.sizeof(abi.Mixture)
    assert out[2] == C.sizeof(abi.StepReport)
    assert out[3] == C.sizeof(abi.ProbeRow)
    assert out[4] == C.sizeof(abi.BatchConfig)
    assert out[5] == C.sizeof(abi.SlotInfo)


def test_presets_validate(lesgpu, abi):
    L = lesgpu.lib()
    for preset in range(5):
        p = lesgpu.params(preset)
        assert L.les_params_validate(C.byref(p)) == 0
    p = lesgpu.params(abi.LES_PRESET_REFERENCE, nx=2)
    assert L.les_params_validate(C.byref(p)) == abi.LES_ERR_PARAM
    p = lesgpu.params(abi.LES_PRESET_REFERENCE, precision=abi.LES_PREC_FP32)
    assert L.les_params_validate(C.byref(p)) == abi.LES_ERR_PARAM


def test_mixture_file_roundtrip(lesgpu, abi, tmp_path):
    L = lesgpu.lib()
    m = lesgpu.mixture()
    chem, tran = str(tmp_path / "chem.bin").encode(), str(tmp_path / "tran.bin").encode()
    assert L.les_mixture_save(C.byref(m), chem, tran) == 0
    m2 = abi.Mixture()
    assert L.les_mixture_load(C.byref(m2), chem, tran) == 0
    assert bytes(m) == bytes(m2)
    assert L.les_mixture_load(C.byref(m2), b"/nonexistent/chem.bin", tran) == abi.LES_ERR_IO


def test_create_without_gpu_fails_loudly(lesgpu):
    """On a machine without a GPU the engine refuses to create a context
    (there is no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(lesgpu.LesError):
        lesgpu.LesGpu()


def test_pool_bytes_counts_probe_rings(lesgpu):
    """les_batch_pool_bytes includes the flame-region probe rings: 8 bytes x
    slots x probe_cap x 16 stations x (2 + 6 ny_max) values."""
    base = dict(max_slots=8, pool_cells=8 * 64 * 48 * 48, max_nx=64, max_ny=48, max_nz=48, precision=1)
    b0 = lesgpu.pool_bytes(lesgpu.batch_config(probe_cap=0, **base))
    b1 = lesgpu.pool_bytes(lesgpu.batch_config(probe_cap=64, **base))
    assert b1 - b0 == 8 * 8 * 64 * 16 * (2 + 6 * 48)
