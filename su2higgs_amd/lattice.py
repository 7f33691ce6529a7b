"""Python handle over libsu2hip.so: a batch of lattices on one GPU stream.

Every method is one C ABI call (include/su2hip.h); arrays are numpy in the
reference's host layout (double[V][3][4] links, double[V][4] doublet, ...).
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import NSUMS, Counters, Muca, Params, check, ptr, volume


class Batch:
    def __init__(self, params, max_slots=None, hbm_budget=0):
        if isinstance(params, Params):
            params = [params]
        self.lib = abi.load()
        n = len(params)
        arr = (Params * n)(*params)
        h = C.c_void_p()
        check(self.lib.su2hip_create(n, arr, max_slots or n, hbm_budget, C.byref(h)))
        self.h = h
        self.params = {i: params[i] for i in range(n)}
        self.max_slots = max_slots or n

    def close(self):
        if getattr(self, "h", None):
            self.lib.su2hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # slots ----------------------------------------------------------------
    def add(self, p):
        s = C.c_int()
        check(self.lib.su2hip_add(self.h, C.byref(p), C.byref(s)))
        self.params[s.value] = p
        return s.value

    def remove(self, slot):
        check(self.lib.su2hip_remove(self.h, slot))
        self.params.pop(slot, None)

    # fields ---------------------------------------------------------------
    def set_fields(self, slot, f):
        """Upload the fields this lattice has (others in f are ignored)."""
        have = abi.field_arrays(self.params[slot])
        check(self.lib.su2hip_set_fields(self.h, slot, *[
            ptr(np.ascontiguousarray(f[k], dtype=np.float64)) if have[k] is not None and f.get(k) is not None
            else None for k in ("su2link", "u1link", "doublet", "triplet", "singlet")]))

    def get_fields(self, slot):
        f = abi.field_arrays(self.params[slot])
        check(self.lib.su2hip_get_fields(self.h, slot, *[ptr(f[k]) for k in
                                                         ("su2link", "u1link", "doublet", "triplet", "singlet")]))
        return f

    def cold_start(self, slot, phi0=0.0, sigma0=0.0, singlet0=0.0):
        check(self.lib.su2hip_cold_start(self.h, slot, phi0, sigma0, singlet0))

    def hot_start(self, slot):
        check(self.lib.su2hip_hot_start(self.h, slot))

    # sweeps ---------------------------------------------------------------
    def update(self, n=1):
        check(self.lib.su2hip_update(self.h, n))

    def pass_(self, kind, parity, d=0):
        check(self.lib.su2hip_pass(self.h, kind, parity, d))

    def sync(self):
        check(self.lib.su2hip_sync(self.h))

    def use_graph(self, on=True):
        check(self.lib.su2hip_use_graph(self.h, 1 if on else 0))

    def set_random_order(self, on=True):
        check(self.lib.su2hip_set_random_order(self.h, 1 if on else 0))

    def muca_acceptance(self, slot, on):
        check(self.lib.su2hip_muca_acceptance(self.h, slot, 1 if on else 0))

    def muca_orderparam(self, slot, recompute=True):
        """(param_value[EVEN], param_value[ODD]) after an optional recompute"""
        pv = np.zeros(2)
        v = C.c_double()
        check(self.lib.su2hip_muca_orderparam(self.h, slot, 1 if recompute else 0,
                                              pv.ctypes.data_as(C.POINTER(C.c_double)), C.byref(v)))
        return pv

    def muca_set_orderparam(self, slot, pv):
        pv = np.ascontiguousarray(pv, dtype=np.float64)
        check(self.lib.su2hip_muca_set_orderparam(self.h, slot, pv.ctypes.data_as(C.POINTER(C.c_double))))

    def set_budget(self, slot, n):
        check(self.lib.su2hip_set_budget(self.h, slot, n))

    def budgets(self):
        b = (C.c_int64 * self.max_slots)()
        check(self.lib.su2hip_get_budgets(self.h, b))
        return list(b)

    def counters(self, slot):
        c = Counters()
        check(self.lib.su2hip_get_counters(self.h, slot, C.byref(c)))
        return c.asdict()

    def reset_counters(self, slot):
        check(self.lib.su2hip_reset_counters(self.h, slot))

    def rng_get(self, slot):
        s, q = C.c_uint64(), C.c_uint64()
        check(self.lib.su2hip_rng_get(self.h, slot, C.byref(s), C.byref(q)))
        return s.value, q.value

    def rng_set(self, slot, seed, pas):
        check(self.lib.su2hip_rng_set(self.h, slot, seed, pas))

    # measurement --------------------------------------------------------------
    def measure(self, slot):
        s = np.zeros(NSUMS)
        check(self.lib.su2hip_measure(self.h, slot, ptr(s)))
        return s

    def measure_all(self):
        s = np.zeros((self.max_slots, NSUMS))
        check(self.lib.su2hip_measure_all(self.h, ptr(s)))
        return s

    # multicanonical -----------------------------------------------------------
    def muca_set(self, slot, m, pos, W):
        check(self.lib.su2hip_muca_set(self.h, slot, C.byref(m),
                                       ptr(np.ascontiguousarray(pos, dtype=np.float64)),
                                       ptr(np.ascontiguousarray(W, dtype=np.float64))))

    def muca_get(self, slot):
        m = Muca()
        check(self.lib.su2hip_muca_get(self.h, slot, C.byref(m), None, None, None))
        n = m.bins + 2 if m.enabled else 0
        pos, W, hits = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
        if n:
            check(self.lib.su2hip_muca_get(self.h, slot, C.byref(m), ptr(pos), ptr(W),
                                           hits.ctypes.data_as(C.POINTER(C.c_int32))))
        return m, pos, W, hits

    # flow ---------------------------------------------------------------------
    def flow(self, slot, scheme, dt, tmax, meas, max_rows=100000):
        rows = np.zeros((max_rows, 1 + NSUMS))
        n = C.c_int()
        check(self.lib.su2hip_flow(self.h, slot, scheme, dt, tmax, meas, ptr(rows), max_rows, C.byref(n)))
        return rows[:min(n.value, max_rows)]

    def flow_start(self, slot, scheme, dt, tmax, meas):
        check(self.lib.su2hip_flow_start(self.h, slot, scheme, dt, tmax, meas))

    def flow_poll(self, slot):
        d = C.c_int()
        check(self.lib.su2hip_flow_poll(self.h, slot, C.byref(d)))
        return bool(d.value)

    def flow_finish(self, slot, max_rows=100000):
        rows = np.zeros((max_rows, 1 + NSUMS))
        n = C.c_int()
        check(self.lib.su2hip_flow_finish(self.h, slot, ptr(rows), max_rows, C.byref(n)))
        return rows[:min(n.value, max_rows)]

    def flow_force(self, slot):
        F = np.zeros((volume(self.params[slot]), 3, 4))
        check(self.lib.su2hip_flow_force(self.h, slot, ptr(F)))
        return F

    # timing -------------------------------------------------------------------
    def set_timing(self, on=True):
        check(self.lib.su2hip_set_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        ms = np.zeros(abi.NKINDS)
        n = (C.c_int64 * abi.NKINDS)()
        check(self.lib.su2hip_kernel_times(self.h, ptr(ms), n))
        return ms, np.array(list(n))

    def reset_timing(self):
        check(self.lib.su2hip_reset_timing(self.h))

    def set_chunk(self, kind, chunk):
        check(self.lib.su2hip_set_chunk(self.h, kind, chunk))

    def chunk(self, kind):
        c = C.c_int()
        check(self.lib.su2hip_get_chunk(self.h, kind, C.byref(c)))
        return c.value

    def set_waves(self, kind, waves):
        check(self.lib.su2hip_set_waves(self.h, kind, waves))

    def waves(self, kind):
        w = C.c_int()
        check(self.lib.su2hip_get_waves(self.h, kind, C.byref(w)))
        return w.value

    def set_hop_cache(self, on=True):
        check(self.lib.su2hip_set_hop_cache(self.h, 1 if on else 0))

    def hop_cache(self):
        v = C.c_int()
        check(self.lib.su2hip_get_hop_cache(self.h, C.byref(v)))
        return bool(v.value)

    def autotune(self, repeats=50):
        """Pick the fastest workgroup chunk (and, where a kind has two, the
        faster occupancy target) per pass kind; returns {kind name: (ms at
        64, 128, 256)} for the kinds of the sweep."""
        t = np.full(abi.NKINDS * 3, np.nan)
        check(self.lib.su2hip_autotune(self.h, repeats, ptr(t)))
        return {abi.KIND_NAMES[k]: tuple(t[3 * k:3 * k + 3]) for k in range(abi.NKINDS)
                if not np.isnan(t[3 * k])}

    def correlators(self, slot, d=-1):
        L = max(self.params[slot].L)
        out = np.zeros((L // 2, 10))
        n = C.c_int()
        check(self.lib.su2hip_correlators(self.h, slot, d, ptr(out), L // 2, C.byref(n)))
        return out[:n.value]

    def blocked_correlators(self, slot, levels):
        L = max(self.params[slot].L)
        nl, nd = C.c_int(), C.c_int()
        buf = np.zeros(max(levels, 1) * (L // 2 + 1) * 10)
        check(self.lib.su2hip_blocked_correlators(self.h, slot, levels, ptr(buf), C.byref(nl), C.byref(nd)))
        return buf[:nl.value * nd.value * 10].reshape(nl.value, nd.value, 10)

    def zprofile(self, slot, d=-1):
        L = max(self.params[slot].L)
        out = np.zeros((L, 4))
        n = C.c_int()
        check(self.lib.su2hip_zprofile(self.h, slot, d, ptr(out), L, C.byref(n)))
        return out[:n.value]

    def wall_start(self, slot, d=-1):
        check(self.lib.su2hip_wall_start(self.h, slot, d))

    def muca_sums(self, slot):
        n = self.muca_get(slot)[0].bins + 2
        g, s = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
        check(self.lib.su2hip_muca_get_sums(self.h, slot, g.ctypes.data_as(C.POINTER(C.c_int64)),
                                            s.ctypes.data_as(C.POINTER(C.c_int64))))
        return g, s

    def muca_set_sums(self, slot, g, s):
        g = np.ascontiguousarray(g, dtype=np.int64)
        s = np.ascontiguousarray(s, dtype=np.int64)
        check(self.lib.su2hip_muca_set_sums(self.h, slot, g.ctypes.data_as(C.POINTER(C.c_int64)),
                                            s.ctypes.data_as(C.POINTER(C.c_int64))))

    def time_pass(self, kind, repeats=100):
        """(average kernel ms, average ms of the back-to-back span) of
        `repeats` launches of one pass kind."""
        k, s = C.c_double(), C.c_double()
        check(self.lib.su2hip_time_pass(self.h, kind, repeats, C.byref(k), C.byref(s)))
        return k.value, s.value


def copy_probe(nbytes, repeats=20):
    lib = abi.load()
    ms = C.c_double()
    check(lib.su2hip_copy_probe(nbytes, repeats, C.byref(ms)))
    return ms.value
