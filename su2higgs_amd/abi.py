"""ctypes mirror of include/su2hip.h (the C ABI of libsu2hip.so).

The structures here are byte-for-byte the C structs; `load()` opens the
in-tree library and raises if it was not built (there is no CPU fallback:
the sweep only runs as HIP kernels).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
LIBPATH = os.path.join(LIBDIR, "libsu2hip.so")

METROPOLIS, HEATBATH, OVERRELAX = 1, 2, 3
EVEN, ODD = 0, 1
PHISQ, SIGMASQ, PHI2MINUSSIGMA2, PHI2SQ = 1, 2, 3, 4
READONLY, FAST, SLOW = 0, 1, 2
FP64, FP32 = 64, 32
OK, EINVAL, ENOMEM, EDEVICE, ERANGE, EUNSUPPORTED, ESTATE, EIO = 0, -1, -2, -3, -4, -5, -6, -7
LINKS_AUTO, LINKS_FULL, LINKS_PACKED = 0, 1, 2
FLOW_EULER, FLOW_RK3 = 1, 3

(K_LINK_HEATBATH, K_LINK_METROPOLIS, K_LINK_OVERRELAX, K_U1_METROPOLIS,
 K_DOUBLET_OVERRELAX, K_DOUBLET_METROPOLIS, K_TRIPLET_OVERRELAX,
 K_TRIPLET_METROPOLIS, K_SINGLET_OVERRELAX, K_SINGLET_METROPOLIS, K_MEASURE,
 K_MUCA, K_FLOW, K_AUX, NKINDS) = range(15)

KIND_NAMES = ["link_heatbath", "link_metropolis", "link_overrelax",
              "u1_metropolis", "doublet_overrelax", "doublet_metropolis",
              "triplet_overrelax", "triplet_metropolis", "singlet_overrelax",
              "singlet_metropolis", "measure", "muca", "flow", "aux"]

SUM_NAMES = ["action", "wilson", "u1wilson", "cov_phi", "phi2", "phi4",
             "hop_sigma", "sigma2", "sigma4", "phi2sigma2", "magcharge",
             "magcharge_abs", "s_cov", "s", "s2", "s3", "s4", "sphisq",
             "s2phisq", "cov_phi_2", "phi2_2", "phi4_2", "r12", "i12"]
NSUMS = len(SUM_NAMES)
SUM = {n: i for i, n in enumerate(SUM_NAMES)}


class Params(C.Structure):
    _fields_ = [
        ("L", C.c_int32 * 3),
        ("nhiggs", C.c_int32), ("u1", C.c_int32), ("triplet", C.c_int32),
        ("singlet", C.c_int32), ("precision", C.c_int32),
        ("algorithm_su2link", C.c_int32), ("algorithm_u1link", C.c_int32),
        ("algorithm_su2doublet", C.c_int32),
        ("algorithm_su2triplet", C.c_int32), ("algorithm_singlet", C.c_int32),
        ("update_links", C.c_int32), ("link_overrelax", C.c_int32),
        ("scalar_sweeps", C.c_int32), ("update_su2doublet", C.c_int32),
        ("update_su2triplet", C.c_int32), ("update_singlet", C.c_int32),
        ("reunitarize", C.c_int32), ("link_format", C.c_int32),
        ("betasu2", C.c_double), ("betau1", C.c_double), ("r_u1", C.c_double),
        ("msq_phi", C.c_double), ("lambda_phi", C.c_double),
        ("msq_triplet", C.c_double), ("a2", C.c_double), ("b4", C.c_double),
        ("b1_s", C.c_double), ("msq_s", C.c_double), ("b3_s", C.c_double),
        ("b4_s", C.c_double), ("a1_s", C.c_double), ("a2_s", C.c_double),
        ("seed", C.c_uint64),
        ("msq_phi2", C.c_double), ("m12sq_re", C.c_double), ("m12sq_im", C.c_double),
        ("lam2", C.c_double), ("lam3", C.c_double), ("lam4", C.c_double),
        ("lam5_re", C.c_double), ("lam5_im", C.c_double), ("lam6_re", C.c_double),
        ("lam6_im", C.c_double), ("lam7_re", C.c_double), ("lam7_im", C.c_double),
    ]


class Counters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "total_su2link", "accepted_su2link", "total_u1link",
        "accepted_u1link", "total_doublet", "accepted_doublet",
        "total_overrelax_doublet", "acc_overrelax_doublet", "total_triplet",
        "accepted_triplet", "total_overrelax_triplet", "acc_overrelax_triplet",
        "total_singlet", "accepted_singlet", "total_overrelax_singlet",
        "acc_overrelax_singlet", "total_muca", "accepted_muca",
        "total_link_or", "accepted_link_or", "total_doublet2",
        "accepted_doublet2", "total_overrelax_doublet2", "acc_overrelax_doublet2")]

    def asdict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Muca(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32), ("orderparam", C.c_int32),
        ("mode", C.c_int32), ("checks_per_sweep", C.c_int32),
        ("bins", C.c_int32), ("last_max", C.c_int32),
        ("do_acceptance", C.c_int32), ("muca_count", C.c_int32),
        ("update_interval", C.c_int32), ("reserved", C.c_int32),
        ("delta", C.c_double), ("min", C.c_double), ("max", C.c_double),
        ("wrk_min", C.c_double), ("wrk_max", C.c_double),
        ("param_value", C.c_double * 2),
    ]


class RunConfig(C.Structure):
    """su2hip_run_config of include/su2hip_io.h."""
    _fields_ = [
        ("p", Params), ("gradflow", C.c_int32), ("reset", C.c_int32),
        ("run_checks", C.c_int32), ("multicanonical", C.c_int32),
        ("measure_local", C.c_int32), ("iterations", C.c_int64),
        ("interval", C.c_int64), ("checkpoint", C.c_int64),
        ("n_thermalize", C.c_int64), ("phi0", C.c_double),
        ("sigma0", C.c_double), ("singlet0", C.c_double),
        ("hot_start", C.c_int32), ("do_flow", C.c_int32),
        ("flow_interval", C.c_int32), ("flow_meas_interval", C.c_int32),
        ("flow_scheme", C.c_int32), ("flow_dt", C.c_double),
        ("flow_t_max", C.c_double), ("muca", Muca),
        ("resultsfile", C.c_char * 200), ("latticefile", C.c_char * 200),
        ("weightfile", C.c_char * 200), ("correlators", C.c_int32),
        ("do_correlators", C.c_int32), ("correlator_interval", C.c_int32),
        ("measure_z", C.c_int32), ("do_z_meas", C.c_int32), ("setup_wall", C.c_int32),
        ("interval_z", C.c_int64), ("blocking", C.c_int32), ("blocks", C.c_int32),
        ("trajectory", C.c_int32), ("do_trajectory", C.c_int32),
    ]


class TrajectoryConfig(C.Structure):
    """su2hip_trajectory_config (include/su2hip_traj.h)."""
    _fields_ = [
        ("n_traj", C.c_int32), ("interval", C.c_int32), ("mode_interval", C.c_int32),
        ("reserved", C.c_int32), ("min", C.c_double), ("max", C.c_double),
    ]


class Request(C.Structure):
    """su2hip_request (include/su2hip_serve.h)."""
    _fields_ = [
        ("id", C.c_int64), ("chain", C.c_int64), ("sweeps", C.c_int64),
        ("arrival", C.c_double), ("betasu2", C.c_double), ("msq_phi", C.c_double),
        ("lambda_phi", C.c_double), ("seed", C.c_uint64), ("L", C.c_int32),
        ("flow", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("id", C.c_int64), ("status", C.c_int32), ("slot", C.c_int32),
        ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("sums", C.c_double * NSUMS), ("flow_t", C.c_double),
        ("flow_sums", C.c_double * NSUMS), ("pass_", C.c_uint64), ("sweeps_done", C.c_int64),
    ]


class ServeConfig(C.Structure):
    _fields_ = [
        ("tmpl", Params), ("hbm_budget", C.c_uint64), ("max_slots", C.c_int32),
        ("tick_sweeps", C.c_int32), ("use_graph", C.c_int32), ("realtime", C.c_int32),
        ("keep_chains", C.c_int32), ("flow_scheme", C.c_int32), ("flow_dt", C.c_double),
        ("flow_t_max", C.c_double), ("phi0", C.c_double), ("sigma0", C.c_double),
        ("singlet0", C.c_double), ("backfill", C.c_int32), ("pad", C.c_int32),
        ("max_head_wait", C.c_double),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("completed", C.c_int64), ("failed", C.c_int64), ("makespan", C.c_double),
        ("p50", C.c_double), ("p99", C.c_double), ("mean", C.c_double),
        ("max_wait", C.c_double), ("site_updates", C.c_double),
        ("site_updates_per_s", C.c_double), ("busy", C.c_double), ("ticks", C.c_int64),
        ("peak_resident", C.c_int64), ("peak_bytes", C.c_uint64), ("flows", C.c_int64),
        ("evictions", C.c_int64),
    ]

    def asdict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def make_params(L=(8, 8, 8), nhiggs=1, u1=0, triplet=0, singlet=0,
                precision=FP64, **kw):
    """Params with the reference's sample-config defaults for the rest."""
    p = Params()
    p.L[:] = list(L)
    p.nhiggs, p.u1, p.triplet, p.singlet = nhiggs, u1, triplet, singlet
    p.precision = precision
    p.algorithm_su2link = HEATBATH
    p.algorithm_u1link = METROPOLIS
    p.algorithm_su2doublet = OVERRELAX
    p.algorithm_su2triplet = OVERRELAX
    p.algorithm_singlet = OVERRELAX
    p.update_links = 1
    p.link_overrelax = 0
    p.scalar_sweeps = 4
    p.update_su2doublet = 1
    p.update_su2triplet = 1
    p.update_singlet = 1
    p.reunitarize = 1 if precision == FP32 else 0
    p.betasu2 = 8.0
    p.betau1 = 8.0
    p.r_u1 = 1.0
    p.msq_phi = -0.25
    p.lambda_phi = 0.02
    p.msq_triplet = -0.5
    p.a2 = 2.7
    p.b4 = 0.12
    p.b1_s, p.msq_s, p.b3_s, p.b4_s, p.a1_s, p.a2_s = 0.3, -0.1, 0.2, 0.3, 0.3, 0.25
    p.seed = 1234
    # two-doublet couplings (used when nhiggs == 2): every term of the
    # potential switched on
    p.msq_phi2, p.m12sq_re, p.m12sq_im = -0.15, 0.02, 0.01
    p.lam2, p.lam3, p.lam4 = 0.03, 0.015, -0.01
    p.lam5_re, p.lam5_im, p.lam6_re, p.lam6_im, p.lam7_re, p.lam7_im = \
        0.008, 0.004, 0.005, -0.003, 0.004, 0.002
    for k, v in kw.items():
        if k == "L":
            continue
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def volume(p):
    return int(p.L[0]) * int(p.L[1]) * int(p.L[2])


def field_arrays(p):
    """Zeroed host arrays in the reference layout for the fields of p."""
    V = volume(p)
    return {
        "su2link": np.zeros((V, 3, 4)),
        "u1link": np.zeros((V, 3)) if p.u1 else None,
        "doublet": (np.zeros((2, V, 4)) if p.nhiggs == 2 else np.zeros((V, 4)))
        if p.nhiggs else None,
        "triplet": np.zeros((V, 3)) if p.triplet else None,
        "singlet": np.zeros(V) if p.singlet else None,
    }


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


_lib = None


def load():
    """Open the in-tree libsu2hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            "libsu2hip.so is not built (%s); run __graft_entry__.build()" % LIBPATH)
    lib = C.CDLL(LIBPATH)
    P = C.POINTER
    dp = P(C.c_double)
    h = C.c_void_p
    sig = {
        "su2hip_error_string": (C.c_char_p, [C.c_int]),
        "su2hip_last_error": (C.c_char_p, []),
        "su2hip_create": (C.c_int, [C.c_int, P(Params), C.c_int, C.c_uint64, P(h)]),
        "su2hip_destroy": (C.c_int, [h]),
        "su2hip_footprint": (C.c_int, [P(Params), C.c_int, P(C.c_uint64)]),
        "su2hip_add": (C.c_int, [h, P(Params), P(C.c_int)]),
        "su2hip_remove": (C.c_int, [h, C.c_int]),
        "su2hip_slots": (C.c_int, [h, P(C.c_int), P(C.c_int)]),
        "su2hip_slot_active": (C.c_int, [h, C.c_int, P(C.c_int)]),
        "su2hip_set_fields": (C.c_int, [h, C.c_int, dp, dp, dp, dp, dp]),
        "su2hip_get_fields": (C.c_int, [h, C.c_int, dp, dp, dp, dp, dp]),
        "su2hip_cold_start": (C.c_int, [h, C.c_int, C.c_double, C.c_double, C.c_double]),
        "su2hip_hot_start": (C.c_int, [h, C.c_int]),
        "su2hip_update": (C.c_int, [h, C.c_int]),
        "su2hip_set_budget": (C.c_int, [h, C.c_int, C.c_int64]),
        "su2hip_get_budgets": (C.c_int, [h, P(C.c_int64)]),
        "su2hip_pass": (C.c_int, [h, C.c_int, C.c_int, C.c_int]),
        "su2hip_use_graph": (C.c_int, [h, C.c_int]),
        "su2hip_set_random_order": (C.c_int, [h, C.c_int]),
        "su2hip_mpi_slices": (C.c_int, [P(C.c_int32), C.c_int, P(C.c_int32), P(C.c_int32)]),
        "su2hip_mpi_site_order": (C.c_int, [P(C.c_int32), C.c_int, P(C.c_int64)]),
        "su2hip_get_random_order": (C.c_int, [h, P(C.c_int)]),
        "su2hip_slot_params": (C.c_int, [h, C.c_int, P(Params)]),
        "su2hip_realtime_config_read": (C.c_int, [C.c_char_p, P(TrajectoryConfig)]),
        "su2hip_trajectory_header": (C.c_int, [C.c_char_p, P(TrajectoryConfig)]),
        "su2hip_trajectories": (C.c_int, [h, C.c_int, P(TrajectoryConfig), C.c_int, C.c_char_p]),
        "su2hip_muca_acceptance": (C.c_int, [h, C.c_int, C.c_int]),
        "su2hip_muca_orderparam": (C.c_int, [h, C.c_int, C.c_int, dp, dp]),
        "su2hip_muca_set_orderparam": (C.c_int, [h, C.c_int, dp]),
        "su2hip_sync": (C.c_int, [h]),
        "su2hip_get_counters": (C.c_int, [h, C.c_int, P(Counters)]),
        "su2hip_reset_counters": (C.c_int, [h, C.c_int]),
        "su2hip_rng_get": (C.c_int, [h, C.c_int, P(C.c_uint64), P(C.c_uint64)]),
        "su2hip_rng_set": (C.c_int, [h, C.c_int, C.c_uint64, C.c_uint64]),
        "su2hip_measure": (C.c_int, [h, C.c_int, dp]),
        "su2hip_measure_noalloc": (C.c_int, [h, C.c_int, dp]),
        "su2hip_measure_all": (C.c_int, [h, dp]),
        "su2hip_measure_columns": (C.c_int, [P(Params), dp, C.c_double, C.c_double, dp, P(C.c_int)]),
        "su2hip_muca_set": (C.c_int, [h, C.c_int, P(Muca), dp, dp]),
        "su2hip_muca_get": (C.c_int, [h, C.c_int, P(Muca), dp, dp, P(C.c_int32)]),
        "su2hip_muca_weight_of": (C.c_int, [h, C.c_int, C.c_double, dp]),
        "su2hip_correlators": (C.c_int, [h, C.c_int, C.c_int, dp, C.c_int, P(C.c_int)]),
        "su2hip_zprofile": (C.c_int, [h, C.c_int, C.c_int, dp, C.c_int, P(C.c_int)]),
        "su2hip_blocked_correlators": (C.c_int, [h, C.c_int, C.c_int, dp, P(C.c_int), P(C.c_int)]),
        "su2hip_wall_start": (C.c_int, [h, C.c_int, C.c_int]),
        "su2hip_muca_set_sums": (C.c_int, [h, C.c_int, P(C.c_int64), P(C.c_int64)]),
        "su2hip_muca_get_sums": (C.c_int, [h, C.c_int, P(C.c_int64), P(C.c_int64)]),
        "su2hip_flow": (C.c_int, [h, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, dp, C.c_int, P(C.c_int)]),
        "su2hip_flow_force": (C.c_int, [h, C.c_int, dp]),
        "su2hip_flow_start": (C.c_int, [h, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]),
        "su2hip_flow_poll": (C.c_int, [h, C.c_int, P(C.c_int)]),
        "su2hip_flow_finish": (C.c_int, [h, C.c_int, dp, C.c_int, P(C.c_int)]),
        "su2hip_set_timing": (C.c_int, [h, C.c_int]),
        "su2hip_kernel_times": (C.c_int, [h, dp, P(C.c_int64)]),
        "su2hip_reset_timing": (C.c_int, [h]),
        "su2hip_copy_probe": (C.c_int, [C.c_uint64, C.c_int, dp]),
        "su2hip_set_chunk": (C.c_int, [h, C.c_int, C.c_int]),
        "su2hip_get_chunk": (C.c_int, [h, C.c_int, P(C.c_int)]),
        "su2hip_set_waves": (C.c_int, [h, C.c_int, C.c_int]),
        "su2hip_get_waves": (C.c_int, [h, C.c_int, P(C.c_int)]),
        "su2hip_set_hop_cache": (C.c_int, [h, C.c_int]),
        "su2hip_get_hop_cache": (C.c_int, [h, P(C.c_int)]),
        "su2hip_autotune": (C.c_int, [h, C.c_int, dp]),
        "su2hip_time_pass": (C.c_int, [h, C.c_int, C.c_int, dp, dp]),
        # host I/O layer (include/su2hip_io.h)
        "su2hip_config_read": (C.c_int, [C.c_char_p, C.c_char_p, P(RunConfig)]),
        "su2hip_config_reread": (C.c_int, [C.c_char_p, P(RunConfig)]),
        "su2hip_lattice_write": (C.c_int, [C.c_char_p, P(Params), C.c_int64, C.c_double, C.c_double,
                                           dp, dp, dp, dp, dp]),
        "su2hip_lattice_read": (C.c_int, [C.c_char_p, P(Params), P(C.c_int64), dp, dp,
                                          dp, dp, dp, dp, dp]),
        "su2hip_weight_load": (C.c_int, [C.c_char_p, P(Muca), dp, dp, C.c_int]),
        "su2hip_weight_save": (C.c_int, [C.c_char_p, P(Muca), dp, dp]),
        "su2hip_weight_params_load": (C.c_int, [C.c_char_p, C.c_int, P(C.c_int64), P(C.c_int64),
                                                P(C.c_int)]),
        "su2hip_weight_params_save": (C.c_int, [C.c_char_p, P(Muca), dp, P(C.c_int64),
                                                P(C.c_int64)]),
        "su2hip_labels_write": (C.c_int, [C.c_char_p, P(Params)]),
        "su2hip_correl_labels_write": (C.c_int, [C.c_char_p, P(Params)]),
        "su2hip_correl_append": (C.c_int, [C.c_char_p, P(Params), C.c_int, dp, C.c_int]),
        "su2hip_zprof_labels_write": (C.c_int, [C.c_char_p, P(Params)]),
        "su2hip_zprof_append": (C.c_int, [C.c_char_p, P(Params), dp, C.c_int]),
        "su2hip_measure_format": (C.c_int, [P(Params), dp, C.c_double, C.c_double, C.c_char_p, C.c_int]),
        "su2hip_print_acceptance": (C.c_int, [C.c_void_p, P(RunConfig), P(Counters)]),
        # request loop (include/su2hip_serve.h)
        "su2hip_serve_defaults": (None, [P(ServeConfig)]),
        "su2hip_serve_updates_per_sweep": (C.c_double, [P(Params)]),
        "su2hip_serve_stream": (C.c_int, [C.c_uint64, C.c_double, C.c_int, C.c_int, P(Request)]),
        "su2hip_serve": (C.c_int, [P(ServeConfig), P(Request), C.c_int, P(Result), P(ServeStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Su2hipError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        lib = load()
        raise Su2hipError("su2hip error %d (%s): %s" % (
            rc, lib.su2hip_error_string(rc).decode(),
            lib.su2hip_last_error().decode()))
    return rc
