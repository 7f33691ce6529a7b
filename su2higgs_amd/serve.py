"""Python handle over the request loop of libsu2hip.so (include/su2hip_serve.h)."""
import ctypes as C

from . import abi
from .abi import Request, Result, ServeConfig, ServeStats, check


def config(**kw):
    """su2hip_serve_defaults() with fields overridden; `tmpl_<name>` keys set
    fields of the parameter template."""
    lib = abi.load()
    c = ServeConfig()
    lib.su2hip_serve_defaults(C.byref(c))
    for k, v in kw.items():
        if k.startswith("tmpl_"):
            setattr(c.tmpl, k[5:], v)
        elif k == "tmpl":
            c.tmpl = v
        else:
            if not hasattr(c, k):
                raise AttributeError(k)
            setattr(c, k, v)
    return c


def stream(seed, rate, n, flow_every=10):
    """The synthetic Poisson request stream (su2hip_serve_stream)."""
    lib = abi.load()
    reqs = (Request * max(n, 1))()
    check(lib.su2hip_serve_stream(seed, rate, n, flow_every, reqs))
    return reqs if n else (Request * 0)()


def requests(rows):
    """Request array from dicts (id defaults to the position, chain to -1)."""
    reqs = (Request * len(rows))()
    for i, r in enumerate(rows):
        reqs[i].id = r.get("id", i)
        reqs[i].chain = r.get("chain", -1)
        for k in ("sweeps", "arrival", "betasu2", "msq_phi", "lambda_phi", "seed", "L", "flow"):
            if k in r:
                setattr(reqs[i], k, r[k])
    return reqs


def serve(cfg, reqs):
    """Run the requests; returns (results, stats dict)."""
    lib = abi.load()
    n = len(reqs)
    res = (Result * max(n, 1))()
    st = ServeStats()
    check(lib.su2hip_serve(C.byref(cfg), reqs, n, res, C.byref(st)))
    return list(res)[:n], st.asdict()


def updates_per_sweep(p):
    return abi.load().su2hip_serve_updates_per_sweep(C.byref(p))
