"""The C-ABI library loads and exports every symbol include/snuqs_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "snuqs_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sq_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("sq_apply", "sq_initialize_basis_z", "sq_program_run", "sq_sample",
              "sq_expect_pauli_sum", "sq_adjoint_gradient", "sq_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "snuqs_amd", "libsnuqs_hip.so"))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string_without_gpu():
    lib = ctypes.CDLL(os.path.join(ROOT, "snuqs_amd", "libsnuqs_hip.so"))
    lib.sq_version.restype = ctypes.c_char_p
    lib.sq_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.sq_version()
    assert isinstance(lib.sq_last_error(), bytes)


def test_tuning_knobs_round_trip_without_gpu(tmp_path):
    """Process-wide launch-shape knobs: validation, and the autotuner's JSON
    save / load applying them (host-side only)."""
    import pytest

    from snuqs_amd import _C, autotune

    keys = ["mfma_blocks_per_cu_c64", "mfma5_pipeline_c128", "mfma5_order_c64"]
    saved = {k: _C.core.tune_get(k) for k in keys}
    try:
        table = {"knobs": {"mfma_blocks_per_cu_c64": 64, "mfma5_pipeline_c128": 0,
                           "mfma5_order_c64": 1}}
        path = str(tmp_path / "tuning.json")
        autotune.save(table, path)
        assert autotune.load(path)["knobs"] == table["knobs"]
        assert [_C.core.tune_get(k) for k in keys] == [64, 0, 1]
        with pytest.raises(ValueError):
            _C.core.tune_set("mfma_blocks_per_cu_c64", 0)
        with pytest.raises(ValueError):
            _C.core.tune_set("no_such_knob", 1)
        for bad in (0, 65):
            with pytest.raises(ValueError):
                _C.core.tune_set("mfma_tiles_per_wave", bad)
        assert _C.core.tune_get("mfma_tiles_per_wave") == 4
    finally:
        for k, v in saved.items():
            _C.core.tune_set(k, v)
