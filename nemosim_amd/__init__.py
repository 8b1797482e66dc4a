"""nemosim_amd -- a HIP (gfx950) simulation backend for the NeMo spiking
neural network simulator, with NeMo's C API (include/nemo.h) and a Python
host interface mirroring NeMo's Python module."""
from .nemo import (Configuration, Network, Simulation, NemoError, run_batch,  # noqa: F401
                   apply_stdp_batch, device_count, eval_libm, nrand_seq, add_plugin_path)
