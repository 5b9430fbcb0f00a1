"""ctypes mirrors of the ABI41 structs, generated from include/ssbl_xpbd_hip.h.

The reference bridge (native_backend.py:21-327) spells the same layouts out by
hand; here they are read from the C header the library is compiled against, so
the Python and C views cannot drift apart.  Sizes are checked against the
values the header pins with static_assert.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

HEADER = Path(__file__).resolve().parent.parent / "include" / "ssbl_xpbd_hip.h"

EXPECTED_SIZES = {
    "SsblXpbdConfig": 224,
    "SsblXpbdMesh": 104,
    "SsblXpbdRuntimeColliders": 56,
    "SsblXpbdForceField": 132,
    "SsblXpbdFrameInputs": 240,
    "SsblXpbdGlobalDynamicSceneInputs": 88,
    "SsblXpbdDiagnostics": 992,
}

_SCALARS = {
    "int": ctypes.c_int,
    "float": ctypes.c_float,
    "long long": ctypes.c_longlong,
    "unsigned int": ctypes.c_uint,
}


def _field_type(ctype: str, structs: dict):
    ctype = ctype.replace("const ", "").strip()
    if ctype.endswith("*"):
        base = ctype[:-1].strip()
        if base in _SCALARS:
            return ctypes.POINTER(_SCALARS[base])
        if base in structs:
            return ctypes.POINTER(structs[base])
        return ctypes.c_void_p
    if ctype in _SCALARS:
        return _SCALARS[ctype]
    if ctype in structs:
        return structs[ctype]
    raise ValueError(f"unsupported field type {ctype!r}")


def _parse(header: Path) -> dict:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    structs: dict = {}
    for match in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
        name, body = match.group(1), match.group(2)
        fields = []
        for line in body.split(";"):
            line = " ".join(line.split())
            if not line:
                continue
            m = re.match(r"(.+?)\s*(\**)\s*(\w+)(\[(\d+)\])?$", line)
            if not m:
                raise ValueError(f"cannot parse field {line!r} of {name}")
            ctype = (m.group(1) + m.group(2)).strip()
            ftype = _field_type(ctype, structs)
            if m.group(5):
                ftype = ftype * int(m.group(5))
            fields.append((m.group(3), ftype))
        structs[name] = type(name, (ctypes.Structure,), {"_fields_": fields})
    for name, size in EXPECTED_SIZES.items():
        got = ctypes.sizeof(structs[name])
        if got != size:
            raise RuntimeError(f"{name} is {got} bytes, the ABI requires {size}")
    return structs


STRUCTS = _parse(HEADER)
Config = STRUCTS["SsblXpbdConfig"]
Mesh = STRUCTS["SsblXpbdMesh"]
RuntimeColliders = STRUCTS["SsblXpbdRuntimeColliders"]
ForceField = STRUCTS["SsblXpbdForceField"]
FrameInputs = STRUCTS["SsblXpbdFrameInputs"]
GlobalDynamicSceneInputs = STRUCTS["SsblXpbdGlobalDynamicSceneInputs"]
Diagnostics = STRUCTS["SsblXpbdDiagnostics"]

CAP_STRETCH_OPTIMIZATION = 1 << 0
CAP_PIN_WEIGHTS = 1 << 1
CAP_GLOBAL_DYNAMIC_SCENE = 1 << 2
CAP_HIP_BATCHED_STEP = 1 << 3

# Entry points declared by the headers, with ctypes signatures.
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_i = ctypes.c_int
ABI41_SIGNATURES = {
    "ssbl_create_solver": (_vp, [ctypes.POINTER(Config), ctypes.POINTER(Mesh)]),
    "ssbl_destroy_solver": (_i, [_vp]),
    "ssbl_reset_solver": (_i, [_vp]),
    "ssbl_update_pin_targets": (_i, [_vp, _ip, _fp, _fp, _i]),
    "ssbl_update_runtime_colliders": (_i, [_vp, ctypes.POINTER(RuntimeColliders)]),
    "ssbl_update_positions": (_i, [_vp, _fp, _i]),
    "ssbl_update_static_triangles": (_i, [_vp, _fp, _i]),
    "ssbl_update_dynamic_triangles": (_i, [_vp, _fp, _i]),
    "ssbl_update_frame_inputs": (_i, [_vp, ctypes.POINTER(FrameInputs)]),
    "ssbl_step_solver": (_i, [_vp, _i, _i]),
    "ssbl_step_solver_ex": (_i, [_vp, _i, _i, _i, _i]),
    "ssbl_download_positions": (_i, [_vp, _fp, _i]),
    "ssbl_get_diagnostics": (_i, [_vp, ctypes.POINTER(Diagnostics)]),
    "ssbl_capabilities": (ctypes.c_uint, []),
    "ssbl_create_global_dynamic_scene": (_vp, []),
    "ssbl_destroy_global_dynamic_scene": (_i, [_vp]),
    "ssbl_update_global_dynamic_scene": (_i, [_vp, ctypes.POINTER(GlobalDynamicSceneInputs)]),
    "ssbl_attach_global_dynamic_scene": (_i, [_vp, _vp, _i, _i, _i]),
    "ssbl_clear_global_dynamic_scene": (_i, [_vp]),
    "ssbl_last_error": (ctypes.c_char_p, []),
}
EXT_SIGNATURES = {
    "ssbl_ext_set_stretch_mode": (_i, [_vp, _i]),
    "ssbl_ext_download_state": (_i, [_vp, _fp, _fp, _fp, _i]),
    "ssbl_ext_self_hash_cells": (_i, [_vp, _ip, _i]),
    "ssbl_ext_self_neighbors": (_i, [_vp, _up, _up, _i]),
    "ssbl_ext_static_sdf": (_i, [_vp, _ip, _fp, _fp, _i]),
    "ssbl_ext_set_static_sdf": (_i, [_vp, _ip, _fp, _fp, _i]),
    "ssbl_ext_create_batch": (_vp, [ctypes.POINTER(_vp), _i]),
    "ssbl_ext_destroy_batch": (_i, [_vp]),
    "ssbl_ext_batch_set_members": (_i, [_vp, ctypes.POINTER(_vp), _i]),
    "ssbl_ext_step_batch": (_i, [_vp, _i, _i, _i]),
    "ssbl_ext_step_batch_ex": (_i, [_vp, _i, _i, _i, _i]),
    "ssbl_ext_time_batch_pass": (_i, [_vp, _i, _fp]),
    "ssbl_ext_time_batch_iteration": (_i, [_vp, _i, _fp]),
    "ssbl_ext_device_bytes": (ctypes.c_longlong, [_vp]),
    "ssbl_ext_global_scene_entries": (ctypes.c_longlong, [_vp]),
    "ssbl_ext_update_global_scene_from_handles": (_i, [_vp, ctypes.POINTER(_vp), _ip, _ip, _fp, _i, ctypes.c_float, ctypes.c_float]),
    "ssbl_ext_synchronize": (_i, []),
    "ssbl_ext_set_graph_mode": (_i, [_i]),
    "ssbl_ext_set_autotune": (_i, [_i]),
    "ssbl_ext_set_batch_parts": (_i, [_i]),
    "ssbl_ext_tuned_pool_limit": (_i, [_vp, _vp]),
    "ssbl_ext_graph_stats": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "ssbl_ext_kernel_timing": (_i, [_i]),
    "ssbl_ext_kernel_stats": (_i, [ctypes.c_char_p, _i]),
    "ssbl_ext_backend_name": (ctypes.c_char_p, []),
    "ssbl_ext_step_solver_ex_noalloc": (_i, [_vp, _i, _i, _i, _i]),
}


def declared_symbols() -> list[str]:
    """Every function name declared in include/*.h."""
    names = []
    for header in sorted(HEADER.parent.glob("*.h")):
        text = header.read_text()
        names += re.findall(r"SSBL_API\s+[\w\s\*]+?\b(ssbl_\w+)\s*\(", text)
    return names


def bind(lib: ctypes.CDLL, with_ext: bool = True) -> ctypes.CDLL:
    sigs = dict(ABI41_SIGNATURES)
    if with_ext:
        sigs.update(EXT_SIGNATURES)
    for name, (restype, argtypes) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib
