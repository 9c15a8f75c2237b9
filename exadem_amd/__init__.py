"""exadem_amd — MI355X (gfx950) implementation of exaDEM's per-timestep hot path.

native      ctypes bindings of include/xdem.h (libxdem_hip.so, the only compute path)
inputs      seeded synthetic workloads (libxdem_host.so generators)
"""
__all__ = ["native", "inputs"]
