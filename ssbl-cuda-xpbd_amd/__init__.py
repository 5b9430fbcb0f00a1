"""MI355X (gfx950) XPBD cloth solver behind the reference's ABI41 C entry points.

Submodules: abi (ctypes mirrors generated from include/ssbl_xpbd_hip.h),
native_backend (the NativeXpbdSolver bridge), scenes (constraint builders and
benchmark scenes), build (hipcc recipe for libssbl_xpbd_hip_abi41.so).
"""
