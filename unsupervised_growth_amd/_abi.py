"""ctypes view of include/hvc_engine.h and the loader of libhvc_engine.so.

The engine is the compiled HIP library built in-tree by
``unsupervised_growth_amd.build``; there is no CPU fallback, so a missing or
unloadable library raises ``EngineNotBuilt``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HVC_ENGINE_LIB", os.path.join(HERE, "libhvc_engine.so"))

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)

HVC_OK, HVC_E_ARG, HVC_E_OOM, HVC_E_HIP, HVC_E_NUMERIC, HVC_E_STATE, HVC_E_CAPACITY, HVC_PENDING = range(8)
STATUS_NAMES = {0: "HVC_OK", 1: "HVC_E_ARG", 2: "HVC_E_OOM", 3: "HVC_E_HIP", 4: "HVC_E_NUMERIC",
                5: "HVC_E_STATE", 6: "HVC_E_CAPACITY", 7: "HVC_PENDING"}

# every entry point declared in include/hvc_engine.h
EXPORTS = ["hvc_create", "hvc_upload_network", "hvc_run_trial", "hvc_download_synapses",
           "hvc_upload_synapses", "hvc_patch_neurons", "hvc_trial_spikes", "hvc_trace_ra", "hvc_trace_i",
           "hvc_kernel_timing", "hvc_set_timing", "hvc_eval_math", "hvc_upload_batch", "hvc_run_batch",
           "hvc_trace_ra_noalloc", "hvc_batch_spikes", "hvc_batch_download_synapses", "hvc_batch_upload_synapses",
           "hvc_pool_footprint",
           "hvc_pool_create", "hvc_pool_admit", "hvc_pool_release", "hvc_pool_run", "hvc_pool_spikes",
           "hvc_pool_download_synapses", "hvc_submit", "hvc_poll", "hvc_get_tuning", "hvc_set_tuning", "hvc_autotune", "hvc_destroy",
           "hvc_last_error", "hvc_set_branch_counting", "hvc_branch_counts"]
BRANCH_COUNTERS = ("activate", "super", "unsuper", "deactivate", "remodel", "ltd_delivery", "ltd_delivery_saturated",
                   "desaturation", "sync_ltp", "sync_ltd", "sync_saturated")
# include/hvc_topology.h
TOPOLOGY_EXPORTS = ["hvc_topology_generate", "hvc_topology_prepare", "hvc_topology_free", "hvc_mt19937"]
# include/hvc_driver.h
DRIVER_EXPORTS = ["hvc_driver_create", "hvc_driver_set_output", "hvc_driver_trial", "hvc_driver_write_state",
                  "hvc_driver_maturation", "hvc_driver_history", "hvc_driver_coordinates", "hvc_driver_replaced",
                  "hvc_driver_connections", "hvc_driver_rr_delays", "hvc_driver_trial_number",
                  "hvc_driver_engine", "hvc_driver_last_error", "hvc_driver_destroy", "hvc_driver_chain_test",
                  "hvc_chain_stats_create", "hvc_chain_stats_add", "hvc_chain_stats_result", "hvc_chain_stats_write",
                  "hvc_chain_stats_destroy", "hvc_activity_update", "hvc_maturation_update"]
ALL_EXPORTS = EXPORTS + TOPOLOGY_EXPORTS + DRIVER_EXPORTS


class EngineNotBuilt(RuntimeError):
    pass


class SynapticParams(ctypes.Structure):
    _fields_ = [("Nss", ctypes.c_int), ("A_P", ctypes.c_double), ("A_D", ctypes.c_double),
                ("T_0", ctypes.c_double), ("T_P", ctypes.c_double), ("T_D", ctypes.c_double),
                ("TAU_P", ctypes.c_double), ("TAU_D", ctypes.c_double), ("BETA", ctypes.c_double),
                ("BETA_SUPERSYNAPSE", ctypes.c_double), ("ACTIVATION_THRESHOLD", ctypes.c_double),
                ("SUPERSYNAPSE_THRESHOLD", ctypes.c_double), ("WEIGHT_MAX", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("ra_model", ctypes.c_int), ("i_model", ctypes.c_int), ("precision", ctypes.c_int),
                ("dt", ctypes.c_double), ("trial_ms", ctypes.c_double), ("sync_ms", ctypes.c_double),
                ("mu_s", ctypes.c_double), ("sigma_s", ctypes.c_double), ("mu_d", ctypes.c_double),
                ("sigma_d", ctypes.c_double), ("ra_poisson", ctypes.c_int), ("i_poisson", ctypes.c_int),
                ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32), ("syn", SynapticParams),
                ("spike_capacity", ctypes.c_int)]


class NetworkDesc(ctypes.Structure):
    _fields_ = [("n_ra", ctypes.c_int), ("n_i", ctypes.c_int),
                ("ra2i_ptr", c_int_p), ("ra2i_tgt", c_int_p), ("ra2i_w", c_double_p), ("ra2i_d", c_double_p),
                ("i2ra_ptr", c_int_p), ("i2ra_tgt", c_int_p), ("i2ra_w", c_double_p), ("i2ra_d", c_double_p),
                ("rr_ptr", c_int_p), ("rr_tgt", c_int_p), ("rr_d", c_double_p), ("rr_w", c_double_p),
                ("act_cnt", c_int_p), ("act_ids", c_int_p), ("sup_cnt", c_int_p), ("sup_ids", c_int_p),
                ("remodeled", c_int_p), ("erest", c_double_p), ("gca", c_double_p),
                ("n_train", ctypes.c_int), ("train_ids", c_int_p), ("spread", c_double_p)]


class Trial(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("training", ctypes.c_int), ("trial_index", ctypes.c_uint32),
                ("first_trial", ctypes.c_int), ("mean_injection_time", ctypes.c_double)]


class TrialOut(ctypes.Structure):
    _fields_ = [("soma_counts", c_int_p), ("soma_times", c_double_p), ("soma_cap", ctypes.c_int),
                ("soma_total", ctypes.c_int),
                ("dend_counts", c_int_p), ("dend_times", c_double_p), ("dend_cap", ctypes.c_int),
                ("dend_total", ctypes.c_int),
                ("i_counts", c_int_p), ("i_times", c_double_p), ("i_cap", ctypes.c_int),
                ("i_total", ctypes.c_int),
                ("bad_values", ctypes.c_int), ("warnings", ctypes.c_int * 3),
                ("ra_state", c_double_p), ("i_state", c_double_p)]


class PoolCaps(ctypes.Structure):
    _fields_ = [("max_nets", ctypes.c_int), ("max_ra", ctypes.c_int), ("max_i", ctypes.c_int),
                ("max_rr", ctypes.c_longlong), ("max_ra2i", ctypes.c_longlong), ("max_i2ra", ctypes.c_longlong),
                ("max_in_degree", ctypes.c_int), ("spikes_per_neuron", ctypes.c_int)]


class Tuning(ctypes.Structure):
    _fields_ = [("integrate_block", ctypes.c_int), ("i_lanes", ctypes.c_int), ("deliver_block", ctypes.c_int),
                ("sync_block", ctypes.c_int), ("ra_lanes", ctypes.c_int), ("deliver_lists", ctypes.c_int)]


class Synapses(ctypes.Structure):
    _fields_ = [("w", c_double_p), ("act_cnt", c_int_p), ("act_ids", c_int_p), ("sup_cnt", c_int_p),
                ("sup_ids", c_int_p), ("remodeled", c_int_p)]


class TopologyParams(ctypes.Structure):
    _fields_ = [("arrangement", ctypes.c_int), ("A_RA2I", ctypes.c_double), ("SIGMA_RA2I", ctypes.c_double),
                ("B_I2RA", ctypes.c_double), ("SIGMA_I2RA", ctypes.c_double), ("Gei_max", ctypes.c_double),
                ("Gie_max", ctypes.c_double), ("delay_constant", ctypes.c_double)]


class Topology(ctypes.Structure):
    _fields_ = [("n_ra", ctypes.c_int), ("n_i", ctypes.c_int), ("n_tr", ctypes.c_int), ("dim", ctypes.c_int),
                ("model_interneuron_distance", ctypes.c_double), ("ra_xyz", c_double_p), ("i_xyz", c_double_p),
                ("ra2i_ptr", c_int_p), ("ra2i_tgt", c_int_p), ("ra2i_w", c_double_p), ("ra2i_len", c_double_p),
                ("ra2i_d", c_double_p),
                ("i2ra_ptr", c_int_p), ("i2ra_tgt", c_int_p), ("i2ra_w", c_double_p), ("i2ra_len", c_double_p),
                ("i2ra_d", c_double_p),
                ("training_clustered", c_int_p), ("training_random", c_int_p)]


class MaturationParamsC(ctypes.Structure):
    _fields_ = [("E_REST_MATURE", ctypes.c_double), ("E_REST_IMMATURE", ctypes.c_double),
                ("GCA_MATURE", ctypes.c_double), ("GCA_IMMATURE", ctypes.c_double),
                ("DEATH_RATE_THRESHOLD", ctypes.c_double), ("RATE_WINDOW_LONG", ctypes.c_int),
                ("ACTIVITY_WINDOW", ctypes.c_int)]


class Geometry(ctypes.Structure):
    _fields_ = [("ra_xyz", c_double_p), ("i_xyz", c_double_p), ("model_interneuron_distance", ctypes.c_double)]


class DriverResume(ctypes.Structure):
    _fields_ = [("mature", c_int_p), ("maturation_scale", c_int_p), ("trials_after_replacement", c_int_p),
                ("history", c_int_p), ("trial_number", ctypes.c_int)]


class DriverReport(ctypes.Structure):
    _fields_ = [("trial", ctypes.c_int), ("mean_injection_time", ctypes.c_double), ("soma_spikes", ctypes.c_int),
                ("dend_spikes", ctypes.c_int), ("i_spikes", ctypes.c_int), ("bad_values", ctypes.c_int),
                ("became_driven", ctypes.c_int), ("matured", ctypes.c_int), ("replaced", ctypes.c_int),
                ("saved", ctypes.c_int)]


_lib = None


def load(path=LIB_PATH):
    """Load libhvc_engine.so (once) and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineNotBuilt("HIP engine library not built: %s (run python -m unsupervised_growth_amd.build)" % path)
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise EngineNotBuilt("cannot load %s: %s" % (path, e))
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.hvc_create.argtypes = [P(Params), ctypes.c_int, P(vp)]
    lib.hvc_upload_network.argtypes = [vp, P(NetworkDesc)]
    lib.hvc_run_trial.argtypes = [vp, P(Trial), P(TrialOut)]
    lib.hvc_download_synapses.argtypes = [vp, P(Synapses)]
    lib.hvc_upload_synapses.argtypes = [vp, P(Synapses)]
    lib.hvc_patch_neurons.argtypes = [vp, c_int_p, ctypes.c_int, c_double_p, c_double_p]
    lib.hvc_trace_ra.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_double_p, c_double_p, ctypes.c_uint32,
                                 c_double_p, c_double_p, c_double_p, c_double_p, c_u8_p]
    lib.hvc_trace_i.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                c_double_p, c_double_p, c_double_p, c_u8_p]
    lib.hvc_trace_ra_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_uint32, vp, vp, vp, vp,
                                         vp]
    lib.hvc_trial_spikes.argtypes = [vp, ctypes.c_int, c_int_p, c_double_p, ctypes.c_int, c_int_p]
    lib.hvc_upload_batch.argtypes = [vp, ctypes.c_int, P(NetworkDesc), P(ctypes.c_uint32)]
    lib.hvc_run_batch.argtypes = [vp, P(Trial), P(TrialOut)]
    lib.hvc_batch_spikes.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_int_p, c_double_p, ctypes.c_int, c_int_p]
    lib.hvc_batch_download_synapses.argtypes = [vp, ctypes.c_int, P(Synapses)]
    lib.hvc_batch_upload_synapses.argtypes = [vp, ctypes.c_int, P(Synapses)]
    ll, sz = ctypes.c_longlong, ctypes.c_size_t
    lib.hvc_pool_footprint.argtypes = [P(Params), P(PoolCaps), ctypes.c_int, ctypes.c_int, ll, ll, ll, P(sz), P(sz)]
    lib.hvc_pool_create.argtypes = [P(Params), ctypes.c_int, P(PoolCaps), P(vp)]
    lib.hvc_pool_admit.argtypes = [vp, P(NetworkDesc), ctypes.c_uint32, ctypes.c_uint32, c_int_p]
    lib.hvc_pool_release.argtypes = [vp, ctypes.c_int]
    lib.hvc_pool_run.argtypes = [vp, ctypes.c_int, c_int_p, P(Trial), P(TrialOut)]
    lib.hvc_pool_spikes.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_int_p, c_double_p, ctypes.c_int, c_int_p]
    lib.hvc_pool_download_synapses.argtypes = [vp, ctypes.c_int, P(Synapses)]
    lib.hvc_submit.argtypes = [vp, ctypes.c_int, P(Trial), P(ctypes.c_uint64)]
    lib.hvc_poll.argtypes = [vp, ctypes.c_uint64, P(TrialOut), P(ll)]
    lib.hvc_get_tuning.argtypes = [vp, P(Tuning)]
    lib.hvc_set_tuning.argtypes = [vp, P(Tuning)]
    lib.hvc_autotune.argtypes = [vp, ctypes.c_int, P(Tuning), c_double_p]
    lib.hvc_kernel_timing.argtypes = [vp, ctypes.c_char_p, c_double_p, c_int_p]
    lib.hvc_set_timing.argtypes = [vp, ctypes.c_int]
    lib.hvc_set_branch_counting.argtypes = [vp, ctypes.c_int]
    lib.hvc_branch_counts.argtypes = [vp, c_int_p]
    lib.hvc_eval_math.argtypes = [vp, ctypes.c_int, ctypes.c_int, c_double_p, c_double_p, c_double_p]
    lib.hvc_destroy.argtypes = [vp]
    lib.hvc_destroy.restype = None
    lib.hvc_last_error.argtypes = [vp]
    lib.hvc_last_error.restype = ctypes.c_char_p
    lib.hvc_topology_generate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(TopologyParams),
                                          ctypes.c_uint32, P(P(Topology))]
    lib.hvc_topology_prepare.argtypes = [P(Topology), P(TopologyParams), ctypes.c_uint32, c_double_p]
    lib.hvc_topology_free.argtypes = [P(Topology)]
    lib.hvc_topology_free.restype = None
    lib.hvc_mt19937.argtypes = [ctypes.c_uint32, ll, ctypes.c_int, P(ctypes.c_uint32)]
    lib.hvc_driver_create.argtypes = [P(Params), ctypes.c_int, P(NetworkDesc), P(Geometry), P(TopologyParams),
                                      P(MaturationParamsC), P(DriverResume), ctypes.c_uint32, P(vp)]
    lib.hvc_driver_set_output.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.hvc_driver_trial.argtypes = [vp, ctypes.c_int, P(DriverReport)]
    lib.hvc_driver_write_state.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.hvc_driver_maturation.argtypes = [vp, c_int_p, c_int_p, c_double_p, c_double_p, c_int_p]
    lib.hvc_driver_history.argtypes = [vp, c_int_p]
    lib.hvc_driver_coordinates.argtypes = [vp, c_double_p, c_double_p]
    lib.hvc_driver_replaced.argtypes = [vp, c_int_p, c_int_p, ctypes.c_int, c_int_p]
    lib.hvc_driver_connections.argtypes = [vp, P(ll), P(ll), c_int_p, c_int_p, c_double_p, c_double_p, c_int_p,
                                           c_int_p, c_double_p, c_double_p]
    lib.hvc_driver_rr_delays.argtypes = [vp, c_double_p]
    lib.hvc_driver_trial_number.argtypes = [vp]
    lib.hvc_driver_engine.argtypes = [vp]
    lib.hvc_driver_engine.restype = vp
    lib.hvc_driver_last_error.argtypes = [vp]
    lib.hvc_driver_last_error.restype = ctypes.c_char_p
    lib.hvc_driver_destroy.argtypes = [vp]
    lib.hvc_driver_destroy.restype = None
    lib.hvc_activity_update.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p,
                                        c_int_p,
                                        c_int_p, ctypes.c_double, c_int_p, c_int_p, c_int_p]
    lib.hvc_maturation_update.argtypes = [ctypes.c_int, P(MaturationParamsC), c_int_p, c_int_p, c_double_p,
                                          c_double_p, c_int_p]
    lib.hvc_driver_chain_test.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, c_double_p]
    lib.hvc_chain_stats_create.argtypes = [ctypes.c_int, P(vp)]
    lib.hvc_chain_stats_add.argtypes = [vp, c_int_p, c_double_p, c_int_p, c_double_p]
    lib.hvc_chain_stats_result.argtypes = [vp, c_int_p, c_double_p]
    lib.hvc_chain_stats_write.argtypes = [vp, ctypes.c_char_p]
    lib.hvc_chain_stats_destroy.argtypes = [vp]
    lib.hvc_chain_stats_destroy.restype = None
    special = ("hvc_destroy", "hvc_last_error", "hvc_topology_free", "hvc_driver_engine", "hvc_driver_last_error",
               "hvc_driver_destroy", "hvc_chain_stats_destroy")
    for name in ALL_EXPORTS:
        if name not in special:
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib
