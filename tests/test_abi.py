"""The C ABI library loads on the CPU and exports every symbol include/bbts_hip.h declares.

No compute calls here: there is no GPU in the build container.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ensure_built():
    import bbts_amd
    if not os.path.exists(bbts_amd.LIB_PATH):
        bbts_amd.build_library()
    return bbts_amd


def _declared():
    text = open(os.path.join(ROOT, "include", "bbts_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bbts_hip_\w+)\s*\(", text)))


def test_header_declares_the_python_binding_set():
    import bbts_amd
    assert _declared() == bbts_amd.declared_symbols()


def test_library_exports_every_declared_symbol():
    bbts_amd = _ensure_built()
    out = subprocess.check_output(["nm", "-D", "--defined-only", bbts_amd.LIB_PATH], text=True)
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    bbts_amd = _ensure_built()
    L = bbts_amd.lib()
    for s in _declared():
        assert getattr(L, s) is not None
    # pure functions that need no device
    assert L.bbts_hip_dense_size(1024, 1024) == 260 + 4 * 1024 * 1024
    assert L.bbts_hip_strerror(2) == b"shape mismatch"
    assert L.bbts_hip_num_tile_configs() >= 4
    assert L.bbts_hip_costmodel_num_features() == 18


def test_costmodel_features_are_host_side():
    _ensure_built()
    from bbts_amd.device import CostModel
    f = CostModel.features(1024, 512, 256, 1, 0, 1)
    assert f.shape == (18,)
    assert abs(f[0] - 10 / 12) < 1e-6 and abs(f[1] - 9 / 12) < 1e-6 and abs(f[2] - 8 / 12) < 1e-6
    assert f[8 + 1] == 1.0 and f[8:14].sum() == 1.0


def test_init_without_gpu_fails_loudly():
    # in the build container there is no device: init must return an error, never fall back
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    bbts_amd = _ensure_built()
    h = ctypes.c_void_p()
    rc = bbts_amd.lib().bbts_hip_init(0, ctypes.byref(h))
    assert rc != 0


def test_runtime_library_exports_every_declared_symbol():
    _ensure_built()
    from bbts_amd import runtime
    text = open(os.path.join(ROOT, "include", "bbts_runtime.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = sorted(set(re.findall(r"\b(bbts_rt_\w+)\s*\(", text)))
    assert declared == runtime.declared_symbols()
    out = subprocess.check_output(["nm", "-D", "--defined-only", runtime.RT_PATH], text=True)
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    assert not [s for s in declared if s not in exported]
    runtime.rtlib()
