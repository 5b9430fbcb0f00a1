"""The C-ABI boundary: struct layouts and exported symbols (no GPU needed)."""
import ctypes
import subprocess

from conftest import HIP_LIB, ORACLE_LIB


def test_struct_sizes_match_reference_layout():
    from ssbl_xpbd_amd import abi
    # sizes probed from the reference header (ssbl_xpbd_cuda.h) with gcc
    for name, size in abi.EXPECTED_SIZES.items():
        assert ctypes.sizeof(abi.STRUCTS[name]) == size, name
    off = {f[0]: getattr(abi.Config, f[0]).offset for f in abi.Config._fields_}
    assert off["dt"] == 36 and off["self_collision"] == 128
    assert abi.FrameInputs.runtime_colliders.offset == 40
    assert abi.FrameInputs.force_fields.offset == 224
    assert abi.Diagnostics.finite_flag.offset == 232


def _exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_hip_library_exports_every_declared_symbol():
    from ssbl_xpbd_amd import abi
    assert HIP_LIB.exists(), "build the HIP library first"
    missing = set(abi.declared_symbols()) - _exported(HIP_LIB)
    assert not missing, missing


def test_hip_library_loads_without_gpu_calls(nb):
    lib = nb.load_library(str(HIP_LIB))
    assert lib.ssbl_capabilities() & 0b10   # PIN_WEIGHTS, required by the reference bridge
    assert lib.ssbl_last_error() == b""


def test_oracle_exports_the_same_symbols():
    from ssbl_xpbd_amd import abi
    missing = set(abi.declared_symbols()) - _exported(ORACLE_LIB)
    assert not missing, missing


def test_invalid_requests_fail_with_message(nb, oracle):
    from ssbl_xpbd_amd import abi
    cfg = abi.Config()
    cfg.dt = 0.0
    assert not oracle.ssbl_create_solver(ctypes.byref(cfg), ctypes.byref(abi.Mesh()))
    assert b"invalid" in oracle.ssbl_last_error()
    assert oracle.ssbl_step_solver_ex(None, 1, 1, 0, 0) == 0
    assert oracle.ssbl_download_positions(None, None, 0) == 0
