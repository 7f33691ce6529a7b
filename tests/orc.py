"""ctypes access to the oracle (oracle/liboracle.so) and to the reference
compiled from source (oracle/_ref/libref_<tag>.so).  Test infrastructure."""
import ctypes as C
import os

import numpy as np

from su2higgs_amd.abi import Counters, Muca, NSUMS, Params, ptr, volume

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REFDIR = os.path.join(ROOT, "oracle", "_ref")

P = C.POINTER
dp = P(C.c_double)
lp = P(C.c_long)
h = C.c_void_p

_orc = None


def oracle_lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: make -C oracle")
        lib = C.CDLL(ORACLE_SO)
        sig = {
            "orc_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
            "orc_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]),
            "orc_create": (h, [P(Params)]),
            "orc_destroy": (None, [h]),
            "orc_volume": (C.c_long, [h]),
            "orc_evensites": (C.c_long, [h]),
            "orc_tables": (None, [h, lp, lp, lp]),
            "orc_set_fields": (None, [h, dp, dp, dp, dp, dp]),
            "orc_get_fields": (None, [h, dp, dp, dp, dp, dp]),
            "orc_cold_start": (None, [h, C.c_double, C.c_double, C.c_double]),
            "orc_hot_start": (None, [h]),
            "orc_rng_set": (None, [h, C.c_uint64, C.c_uint64]),
            "orc_rng_get": (None, [h, P(C.c_uint64), P(C.c_uint64)]),
            "orc_get_counters": (None, [h, P(Counters)]),
            "orc_reset_counters": (None, [h]),
            "orc_link_staple": (None, [h, C.c_long, C.c_int, dp]),
            "orc_wilson_staple": (None, [h, C.c_long, C.c_int, dp]),
            "orc_doublet_staple": (None, [h, C.c_long, dp]),
            "orc_triplet_staple": (None, [h, C.c_long, dp]),
            "orc_localact_su2link": (C.c_double, [h, C.c_long, C.c_int]),
            "orc_localact_u1link": (C.c_double, [h, C.c_long, C.c_int]),
            "orc_localact_doublet": (C.c_double, [h, C.c_long]),
            "orc_doublet_staple_h": (None, [h, C.c_long, C.c_int, dp]),
            "orc_localact_doublet_h": (C.c_double, [h, C.c_long, C.c_int]),
            "orc_localact_triplet": (C.c_double, [h, C.c_long]),
            "orc_localact_singlet": (C.c_double, [h, C.c_long]),
            "orc_action_local": (C.c_double, [h, C.c_long]),
            "orc_polysolve3": (C.c_double, [C.c_double] * 4),
            "orc_site_update": (C.c_int, [h, C.c_int, C.c_long, C.c_int, C.c_uint64]),
            "orc_pass": (None, [h, C.c_int, C.c_int, C.c_int]),
            "orc_update": (None, [h, C.c_int]),
            "orc_set_random_order": (None, [h, C.c_int]),
            "orc_muca_set_acceptance": (None, [h, C.c_int]),
            "orc_word": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]),
            "orc_measure": (None, [h, dp]),
            "orc_muca_set": (None, [h, P(Muca), dp, dp]),
            "orc_muca_get": (None, [h, P(Muca), dp, dp, P(C.c_int32)]),
            "orc_muca_set_sums": (None, [h, P(C.c_int64), P(C.c_int64)]),
            "orc_correlators": (C.c_int, [h, C.c_int, dp]),
            "orc_longest_dir": (C.c_int, [h]),
            "orc_zprofile": (C.c_int, [h, C.c_int, dp]),
            "orc_wall_start": (None, [h, C.c_int]),
            "orc_blocked_correlators": (C.c_int, [h, C.c_int, dp]),
            "orc_muca_get_sums": (None, [h, P(C.c_int64), P(C.c_int64)]),
            "orc_muca_weight": (C.c_double, [h, C.c_double]),
            "orc_muca_whichbin": (C.c_int, [h, C.c_double]),
            "orc_muca_orderparam": (C.c_double, [h, C.c_int]),
            "orc_flow_force": (None, [h, dp, dp]),
            "orc_flow_step": (None, [h, C.c_int, C.c_double]),
            "orc_flow": (C.c_int, [h, C.c_int, C.c_double, C.c_double, C.c_int, dp, C.c_int]),
        }
        for n, (r, a) in sig.items():
            fn = getattr(lib, n)
            fn.restype = r
            fn.argtypes = a
        _orc = lib
    return _orc


def philox(ctr, key):
    lib = oracle_lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def uniform(seed, pas, site, k):
    return oracle_lib().orc_uniform(seed, pas, site, k)


FIELDS = ("su2link", "u1link", "doublet", "triplet", "singlet")


def _arrs(p):
    V = volume(p)
    return {"su2link": np.zeros((V, 3, 4)), "u1link": np.zeros((V, 3)),
            "doublet": np.zeros((2, V, 4)) if p.nhiggs == 2 else np.zeros((V, 4)),
            "triplet": np.zeros((V, 3)),
            "singlet": np.zeros(V)}


class Oracle:
    def __init__(self, p):
        self.lib = oracle_lib()
        self.p = p
        self.h = self.lib.orc_create(C.byref(p))
        self.V = volume(p)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def tables(self):
        V = self.V
        nx = np.zeros((V, 3), dtype=np.int64)
        pv = np.zeros((V, 3), dtype=np.int64)
        nat = np.zeros(V, dtype=np.int64)
        self.lib.orc_tables(self.h, nx.ctypes.data_as(lp), pv.ctypes.data_as(lp),
                            nat.ctypes.data_as(lp))
        return nx, pv, nat

    def set_fields(self, f):
        self.lib.orc_set_fields(self.h, *[ptr(f.get(k)) if f.get(k) is not None else None for k in FIELDS])

    def get_fields(self):
        f = _arrs(self.p)
        self.lib.orc_get_fields(self.h, *[ptr(f[k]) for k in FIELDS])
        return f

    def cold_start(self, phi0=1.0, sigma0=0.2, singlet0=0.2):
        self.lib.orc_cold_start(self.h, phi0, sigma0, singlet0)

    def hot_start(self):
        self.lib.orc_hot_start(self.h)

    def rng_set(self, seed, pas):
        self.lib.orc_rng_set(self.h, seed, pas)

    def rng_get(self):
        s, q = C.c_uint64(), C.c_uint64()
        self.lib.orc_rng_get(self.h, C.byref(s), C.byref(q))
        return s.value, q.value

    def counters(self):
        c = Counters()
        self.lib.orc_get_counters(self.h, C.byref(c))
        return c.asdict()

    def update(self, n=1):
        self.lib.orc_update(self.h, n)

    def pass_(self, kind, parity, d=0):
        self.lib.orc_pass(self.h, kind, parity, d)

    def site(self, kind, i, d, pas):
        return self.lib.orc_site_update(self.h, kind, i, d, pas)

    def measure(self):
        s = np.zeros(NSUMS)
        self.lib.orc_measure(self.h, ptr(s))
        return s

    def vec(self, fn, n, *args):
        out = np.zeros(n)
        getattr(self.lib, fn)(self.h, *args, ptr(out))
        return out

    def muca_set(self, m, pos, W):
        self.lib.orc_muca_set(self.h, C.byref(m), ptr(np.ascontiguousarray(pos, dtype=np.float64)),
                              ptr(np.ascontiguousarray(W, dtype=np.float64)))

    def muca_get(self):
        m = Muca()
        self.lib.orc_muca_get(self.h, C.byref(m), None, None, None)
        n = m.bins + 2
        pos, W, hits = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
        self.lib.orc_muca_get(self.h, C.byref(m), ptr(pos), ptr(W),
                              hits.ctypes.data_as(P(C.c_int32)))
        return m, pos, W, hits

    def correlators(self, d):
        out = np.zeros((self.p.L[d] // 2, 10))
        n = self.lib.orc_correlators(self.h, d, ptr(out))
        return out[:n]

    def zprofile(self, d=-1):
        out = np.zeros((max(self.p.L), 4))
        n = self.lib.orc_zprofile(self.h, d, ptr(out))
        return out[:n]

    def wall_start(self, d=-1):
        self.lib.orc_wall_start(self.h, d)

    def blocked_correlators(self, levels):
        L = self.p.L[self.lib.orc_longest_dir(self.h)]
        out = np.zeros((max(levels, 1), L // 2, 10))
        n = self.lib.orc_blocked_correlators(self.h, levels, ptr(out))
        return out[:n]

    def muca_sums(self):
        n = self.muca_get()[0].bins + 2
        g, s = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
        self.lib.orc_muca_get_sums(self.h, g.ctypes.data_as(P(C.c_int64)), s.ctypes.data_as(P(C.c_int64)))
        return g, s

    def muca_set_sums(self, g, s):
        g = np.ascontiguousarray(g, dtype=np.int64)
        s = np.ascontiguousarray(s, dtype=np.int64)
        self.lib.orc_muca_set_sums(self.h, g.ctypes.data_as(P(C.c_int64)), s.ctypes.data_as(P(C.c_int64)))

    def flow_force(self):
        F = np.zeros((self.V, 3, 4))
        T = np.zeros((self.V, 3))
        self.lib.orc_flow_force(self.h, ptr(F), ptr(T))
        return F, T

    def flow(self, scheme, dt, tmax, meas, max_rows=100000):
        rows = np.zeros((max_rows, 1 + NSUMS))
        n = self.lib.orc_flow(self.h, scheme, dt, tmax, meas, ptr(rows), max_rows)
        return rows[:min(n, max_rows)]


# --------------------------------------------------------------------------
_refs = {}

REF_FLAGS = {1: "h", 3: "hu", 4: "t", 5: "ht", 17: "hg", 20: "tg", 11: "hsu",
             33: "hc", 36: "tc", 37: "htc", 65: "hz", 68: "tz", 75: "hsuz", 164: "tcb",
             257: "d", 259: "du", 275: "dug", 513: "hr", 516: "tr"}
# builds that do not link as shipped (check_set is undefined in
# src/hb_trajectory.c): their harness libraries load with lazy binding
LAZY = ("hr", "tr")


def ref_tag(p, gradflow=False, correlators=False, measure_z=False, blocking=False,
            trajectory=False):
    m = (1 if p.nhiggs else 0) | (2 if p.u1 else 0) | (4 if p.triplet else 0) \
        | (8 if p.singlet else 0) | (16 if gradflow else 0) | (32 if correlators else 0) \
        | (64 if measure_z else 0) | (128 if blocking else 0) | (256 if p.nhiggs == 2 else 0) \
        | (512 if trajectory else 0)
    return REF_FLAGS.get(m)


def ref_available(tag):
    return tag is not None and os.path.exists(os.path.join(REFDIR, "libref_%s.so" % tag))


def ref_lib(tag):
    if tag in _refs:
        return _refs[tag]
    path = os.path.join(REFDIR, "libref_%s.so" % tag)
    lib = C.CDLL(path, mode=os.RTLD_LAZY) if tag in LAZY else C.CDLL(path)
    sig = {
        "ref_model": (C.c_int, []),
        "ref_setup": (C.c_int, [P(Params)]),
        "ref_evensites": (C.c_long, []),
        "ref_tables": (None, [lp, lp]),
        "ref_put_fields": (None, [dp] * 5),
        "ref_get_fields": (None, [dp] * 5),
        "ref_rng_set": (None, [C.c_uint64, C.c_uint64]),
        "ref_rng_pass": (C.c_uint64, []),
        "ref_counters": (None, [P(Counters)]),
        "ref_update": (None, [C.c_int]),
        "ref_site": (C.c_int, [C.c_int, C.c_long, C.c_int, C.c_uint64]),
        "ref_link_staple": (None, [C.c_long, C.c_int, dp]),
        "ref_wilson_staple": (None, [C.c_long, C.c_int, dp]),
        "ref_localact_su2link": (C.c_double, [C.c_long, C.c_int]),
        "ref_action_local": (C.c_double, [C.c_long]),
        "ref_polysolve3": (C.c_double, [C.c_double] * 4),
        "ref_measure_sums": (None, [dp]),
        "ref_measure_text": (C.c_int, [C.c_char_p, C.c_int]),
        "ref_print_labels": (None, []),
        "ref_muca_load": (C.c_int, [C.c_char_p, P(Muca)]),
        "ref_muca_get": (None, [P(Muca), dp, dp, P(C.c_int32)]),
        "ref_muca_set_mode": (None, [C.c_int]),
        "ref_muca_set_interval": (None, [C.c_int]),
        "ref_muca_sums": (None, [P(C.c_int64), P(C.c_int64)]),
        "ref_muca_weight": (C.c_double, [C.c_double]),
        "ref_muca_whichbin": (C.c_int, [C.c_double]),
        "ref_muca_orderparam": (C.c_double, [C.c_int]),
        "ref_save_lattice": (None, [C.c_char_p, C.c_long]),
        "ref_load_lattice": (C.c_long, [C.c_char_p]),
        "ref_get_parameters": (C.c_int, [C.c_char_p, P(Params), lp, P(Muca), C.c_char_p, C.c_char_p]),
        "ref_set_random_sweeps": (None, [C.c_int]),
        "ref_muca_set_acceptance": (None, [C.c_int]),
    }
    opt = {
        "ref_doublet_staple": (None, [C.c_long, dp]),
        "ref_localact_doublet": (C.c_double, [C.c_long]),
        "ref_doublet_staple_h": (None, [C.c_long, C.c_int, dp]),
        "ref_localact_doublet_h": (C.c_double, [C.c_long, C.c_int]),
        "ref_localact_u1link": (C.c_double, [C.c_long, C.c_int]),
        "ref_localact_triplet": (C.c_double, [C.c_long]),
        "ref_magcharge_cube": (C.c_double, [C.c_long]),
        "ref_localact_singlet": (C.c_double, [C.c_long]),
        "ref_flow_force": (None, [dp, dp]),
        "ref_flow_euler_step": (None, [C.c_double]),
        "ref_grad_flow": (None, [C.c_double, C.c_double, C.c_int, C.c_int]),
        "ref_correlators": (C.c_int, [C.c_int, dp]),
        "ref_measure_correlators": (None, [C.c_char_p, C.c_int, C.c_int]),
        "ref_print_labels_correlators": (None, []),
        "ref_longest_dir": (C.c_int, []),
        "ref_print_z_labels": (None, []),
        "ref_measure_along_z": (None, [C.c_long]),
        "ref_z_dir": (C.c_int, []),
        "ref_prepare_wall": (None, []),
        "ref_blocked_setup": (C.c_int, [C.c_int]),
        "ref_blocked_correlators": (None, [C.c_int]),
        "ref_blocked_values": (C.c_int, [dp]),
        "ref_trajectories": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]),
    }
    for n, (r, a) in sig.items():
        fn = getattr(lib, n)
        fn.restype = r
        fn.argtypes = a
    for n, (r, a) in opt.items():
        if hasattr(lib, n):
            fn = getattr(lib, n)
            fn.restype = r
            fn.argtypes = a
    _refs[tag] = lib
    return lib


class Ref:
    """The reference compiled from source, one global lattice per tag."""

    def __init__(self, p, gradflow=False, correlators=False, measure_z=False, blocking=False,
                 trajectory=False):
        self.tag = ref_tag(p, gradflow, correlators or blocking, measure_z, blocking, trajectory)
        self.lib = ref_lib(self.tag)
        self.p = p
        self.V = volume(p)
        self.lib.ref_setup(C.byref(p))

    def set_fields(self, f):
        self.lib.ref_put_fields(*[ptr(f.get(k)) if f.get(k) is not None else None for k in FIELDS])

    def get_fields(self):
        f = _arrs(self.p)
        self.lib.ref_get_fields(*[ptr(f[k]) for k in FIELDS])
        return f

    def tables(self):
        nx = np.zeros((self.V, 3), dtype=np.int64)
        pv = np.zeros((self.V, 3), dtype=np.int64)
        self.lib.ref_tables(nx.ctypes.data_as(lp), pv.ctypes.data_as(lp))
        return nx, pv

    def rng_set(self, seed, pas):
        self.lib.ref_rng_set(seed, pas)

    def counters(self):
        c = Counters()
        self.lib.ref_counters(C.byref(c))
        return c.asdict()

    def update(self, n=1):
        self.lib.ref_update(n)

    def site(self, kind, i, d, pas):
        return self.lib.ref_site(kind, i, d, pas)

    def measure(self):
        s = np.zeros(NSUMS)
        self.lib.ref_measure_sums(ptr(s))
        return s

    def correlators(self, d):
        out = np.zeros((self.p.L[d] // 2, 10))
        n = self.lib.ref_correlators(d, ptr(out))
        return out[:n]

    def measure_text(self):
        buf = C.create_string_buffer(8192)
        self.lib.ref_measure_text(buf, 8192)
        return buf.value.decode()

    def vec(self, fn, n, *args):
        out = np.zeros(n)
        getattr(self.lib, fn)(*args, ptr(out))
        return out
