"""MI355X (gfx950) Monte Carlo playout engine for 8x8 checkers, drop-in behind gpu_ai's
PlayoutDriver::runPlayouts (src/playout.hpp:27-33).

The product is the C ABI in include/gpuai/mi_playout.h (libgpuai_mi355x.so, built in-tree by
gpu_ai_amd.build) and the C++ host layer under gpu_ai_amd/host; gpu_ai_amd.engine binds the C ABI
for Python.
"""
__all__ = ["build", "engine"]
