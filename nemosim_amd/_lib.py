"""Loader of libnemo_hip.so, the HIP backend with NeMo's C API.

The library is built in-tree (``make -C nemosim_amd``); importing this module
without it raises ImportError -- there is no CPU fallback. The environment
variable NEMOSIM_AMD_LIBRARY names another build of the same library (for
example one compiled with different kernel options) to load instead.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NEMOSIM_AMD_LIBRARY") or os.path.join(_HERE, "libnemo_hip.so")

c_uint_p = ctypes.POINTER(ctypes.c_uint)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_u8_p = ctypes.POINTER(ctypes.c_ubyte)
c_ull_p = ctypes.POINTER(ctypes.c_ulonglong)
c_size_p = ctypes.POINTER(ctypes.c_size_t)
c_void_p = ctypes.c_void_p


class ScheduleT(ctypes.Structure):
    _fields_ = [
        ("fstim_ptr", c_uint_p),
        ("fstim_idx", c_uint_p),
        ("istim_ptr", c_uint_p),
        ("istim_idx", c_uint_p),
        ("istim_val", c_float_p),
    ]


class InfoT(ctypes.Structure):
    _fields_ = [
        ("neurons", ctypes.c_uint),
        ("partitions", ctypes.c_uint),
        ("partition_size", ctypes.c_uint),
        ("max_delay", ctypes.c_uint),
        ("saturating", ctypes.c_uint),
        ("stdp", ctypes.c_uint),
        ("synapses", ctypes.c_ulonglong),
        ("plastic_synapses", ctypes.c_ulonglong),
        ("device_bytes", ctypes.c_ulonglong),
        ("lds_bytes", ctypes.c_uint),
        ("block_threads", ctypes.c_uint),
        ("last_kernel_ms", ctypes.c_float),
        ("fused", ctypes.c_uint),
        ("deferred", ctypes.c_uint),
        ("lanes_per_source", ctypes.c_uint),
        ("stdp_staged_chunks", ctypes.c_uint),
        ("stdp_prefix_ring", ctypes.c_uint),
        ("device_built", ctypes.c_uint),
        ("exchange_waves", ctypes.c_uint),
        ("far_list_sources", ctypes.c_uint),
    ]


# name -> (restype, argtypes); every symbol declared in include/nemo.h and
# include/nemo_hip.h
SIGNATURES = {
    "nemo_version": (ctypes.c_char_p, []),
    "nemo_add_plugin_path": (ctypes.c_int, [ctypes.c_char_p]),
    "nemo_cuda_device_count": (ctypes.c_int, [c_uint_p]),
    "nemo_cuda_device_description": (ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(ctypes.c_char_p)]),
    "nemo_new_configuration": (c_void_p, []),
    "nemo_delete_configuration": (None, [c_void_p]),
    "nemo_log_stdout": (ctypes.c_int, [c_void_p]),
    "nemo_set_stdp_function": (ctypes.c_int, [c_void_p, c_float_p, ctypes.c_size_t, c_float_p, ctypes.c_size_t,
                                              ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "nemo_set_cpu_backend": (ctypes.c_int, [c_void_p]),
    "nemo_set_cuda_backend": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "nemo_cuda_device": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "nemo_backend": (ctypes.c_int, [c_void_p, c_uint_p]),
    "nemo_backend_description": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_char_p)]),
    "nemo_set_write_only_synapses": (ctypes.c_int, [c_void_p]),
    "nemo_new_network": (c_void_p, []),
    "nemo_delete_network": (None, [c_void_p]),
    "nemo_add_neuron_type": (ctypes.c_int, [c_void_p, ctypes.c_char_p, c_uint_p]),
    "nemo_add_neuron_iz": (ctypes.c_int, [c_void_p, ctypes.c_uint] + [ctypes.c_float] * 7),
    "nemo_add_neuron": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_add_synapse": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_float,
                                        ctypes.c_ubyte, c_u64_p]),
    "nemo_neuron_count": (ctypes.c_int, [c_void_p, c_uint_p]),
    "nemo_new_simulation": (c_void_p, [c_void_p, c_void_p]),
    "nemo_delete_simulation": (None, [c_void_p]),
    "nemo_step": (ctypes.c_int, [c_void_p, c_uint_p, ctypes.c_size_t, c_uint_p, c_float_p, ctypes.c_size_t,
                                 ctypes.POINTER(c_uint_p), c_size_p]),
    "nemo_apply_stdp": (ctypes.c_int, [c_void_p, ctypes.c_float]),
    "nemo_get_membrane_potential": (ctypes.c_int, [c_void_p, ctypes.c_uint, c_float_p]),
    "nemo_get_neuron_state_n": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_get_neuron_parameter_n": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_get_neuron_state_s": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_get_neuron_parameter_s": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_get_synapse_source_n": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_uint_p]),
    "nemo_get_synapse_source_s": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_uint_p]),
    "nemo_get_synapse_target_n": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_uint_p]),
    "nemo_get_synapse_target_s": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_uint_p]),
    "nemo_get_synapse_delay_n": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_uint_p]),
    "nemo_get_synapse_delay_s": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_uint_p]),
    "nemo_get_synapse_weight_n": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_float_p]),
    "nemo_get_synapse_weight_s": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_float_p]),
    "nemo_get_synapse_plastic_n": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_u8_p]),
    "nemo_get_synapse_plastic_s": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_u8_p]),
    "nemo_get_synapses_from_n": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.POINTER(c_u64_p), c_size_p]),
    "nemo_get_synapses_from_s": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.POINTER(c_u64_p), c_size_p]),
    "nemo_set_neuron_n": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_set_neuron_s": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, c_float_p]),
    "nemo_set_neuron_iz_n": (ctypes.c_int, [c_void_p, ctypes.c_uint] + [ctypes.c_float] * 7),
    "nemo_set_neuron_iz_s": (ctypes.c_int, [c_void_p, ctypes.c_uint] + [ctypes.c_float] * 7),
    "nemo_set_neuron_state_n": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    "nemo_set_neuron_parameter_n": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    "nemo_set_neuron_state_s": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    "nemo_set_neuron_parameter_s": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float]),
    "nemo_elapsed_wallclock": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_ulong)]),
    "nemo_elapsed_simulation": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_ulong)]),
    "nemo_reset_timer": (ctypes.c_int, [c_void_p]),
    "nemo_strerror": (ctypes.c_char_p, []),
    # nemo_hip.h
    "nemo_add_neurons": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.c_size_t, c_uint_p, ctypes.c_uint, c_float_p]),
    "nemo_add_synapses_csr": (ctypes.c_int, [c_void_p, ctypes.c_size_t, c_uint_p, c_u64_p, c_uint_p, c_u8_p,
                                             c_float_p, c_u8_p]),
    "nemo_set_hip_backend": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "nemo_set_partition_size": (ctypes.c_int, [c_void_p, ctypes.c_uint]),
    "nemo_set_delivery_lanes": (ctypes.c_int, [c_void_p, ctypes.c_uint]),
    "nemo_set_stdp_lds_staging": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "nemo_set_exchange_waves": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "nemo_copy_configuration": (c_void_p, [c_void_p]),
    "nemo_disable_logging": (ctypes.c_int, [c_void_p]),
    "nemo_logging_enabled": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "nemo_partition_size": (ctypes.c_int, [c_void_p, c_uint_p]),
    "nemo_write_only_synapses": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "nemo_max_delay": (ctypes.c_int, [c_void_p, c_uint_p]),
    "nemo_min_weight": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "nemo_max_weight": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "nemo_set_device_construction": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "nemo_set_step_mode": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "nemo_hip_run": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.POINTER(ScheduleT), c_ull_p]),
    "nemo_hip_run_static": (ctypes.c_int, [c_void_p, ctypes.c_uint, ctypes.POINTER(ScheduleT), c_void_p,
                                           c_ull_p]),
    "nemo_hip_read_fired": (ctypes.c_int, [c_void_p, c_uint_p, c_uint_p, ctypes.c_size_t, c_size_p]),
    "nemo_hip_read_raster": (ctypes.c_int, [c_void_p, c_u64_p, ctypes.c_size_t, c_size_p]),
    "nemo_hip_run_batch": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.c_size_t, c_uint_p, c_ull_p]),
    "nemo_hip_get_info": (ctypes.c_int, [c_void_p, ctypes.POINTER(InfoT)]),
    "nemo_hip_last_status": (ctypes.c_int, []),
    "nemo_hip_eval_libm": (ctypes.c_int, [ctypes.c_int, c_float_p, c_float_p, ctypes.c_size_t, ctypes.c_int]),
    "nemo_hip_apply_stdp_batch": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.c_size_t, c_float_p]),
    "nemo_hip_nrand_seq": (ctypes.c_int, [c_u32_p, c_float_p, ctypes.c_size_t]),
}

_lib = None


def load():
    """Return the loaded library; raises ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "nemosim_amd: %s is missing -- build it with `make -C nemosim_amd` "
            "(or __graft_entry__.build()); there is no CPU fallback" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
