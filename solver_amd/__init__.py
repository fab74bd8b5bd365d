"""MI355X-native hot path of the voxel PEEC FFT solver (Draco112358/Solver).

solver_amd.peec mirrors the reference's Julia functions on top of the C ABI
in libpeec_hip.so (include/peec_hip.h).  The library must be built first
(python solver_amd/build.py); nothing here falls back to CPU arithmetic.
"""
