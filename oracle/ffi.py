"""ctypes binding of the CPU oracle (oracle/libplis_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's ``cpu_baseline`` leg.  The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplis_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Network(C.Structure):
    _fields_ = [("S", C.c_int32), ("R", C.c_int32), ("C", C.c_int32),
                ("reac", C.POINTER(C.c_int32)), ("prod", C.POINTER(C.c_int32)),
                ("k", C.POINTER(C.c_double)), ("d", C.POINTER(C.c_double))]


class Topology(C.Structure):
    _fields_ = [("N", C.c_uint32), ("E", C.c_uint32),
                ("row_ptr", C.POINTER(C.c_uint32)), ("col", C.POINTER(C.c_uint32)),
                ("w", C.POINTER(C.c_float)), ("comp", C.POINTER(C.c_int32))]


class Params(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("n_critical", C.c_int32), ("seed", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("t", C.c_double), ("steps", C.c_uint64), ("leapbacks", C.c_uint64),
                ("firings", C.c_uint64), ("subvolume_leaps", C.c_uint64),
                ("crit_events", C.c_uint64), ("status", C.c_int32)]


def _declare(L):
    P = C.POINTER
    L.orc_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    for f in ("orc_exp", "orc_log", "orc_sqrt"):
        getattr(L, f).argtypes = [C.c_double]
        getattr(L, f).restype = C.c_double
    L.orc_logfact.argtypes = [C.c_int64]
    L.orc_logfact.restype = C.c_double
    L.orc_poisson_many.argtypes = [C.c_double, C.c_uint64, C.c_uint32, P(C.c_int64)]
    L.orc_poisson.argtypes = [C.c_double, C.c_uint32, P(C.c_uint32), P(C.c_uint32), C.c_uint32]
    L.orc_poisson.restype = C.c_int64
    L.orc_sq_exp_neg.argtypes = [C.c_float]
    L.orc_sq_exp_neg.restype = C.c_float
    L.orc_sq_decide_band.argtypes = [C.c_float, C.c_float, C.c_float]
    L.orc_sq_decide_band.restype = C.c_int
    L.orc_sq_poisson.argtypes = [C.c_double, C.c_uint32]
    L.orc_sq_poisson.restype = C.c_int
    L.orc_sq_maxerr.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_sq_maxerr.restype = C.c_double
    L.orc_sq_hop_check.argtypes = [P(C.c_double), P(C.c_double), P(C.c_uint32), P(C.c_float),
                                   P(C.c_uint32), C.c_int32, P(C.c_int32), P(C.c_int64)]
    L.orc_sq_hop_check.restype = None
    L.orc_tau_all.argtypes = [P(Network), P(Topology), P(Params), P(C.c_uint32), C.c_uint32,
                              C.c_uint32, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.orc_critical_choice.argtypes = [P(Network), P(Topology), P(Params), P(C.c_uint32), C.c_uint32,
                                      C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_int32),
                                      P(C.c_int32), P(C.c_uint32)]
    L.orc_run.argtypes = [P(Network), P(Topology), P(Params), C.c_uint32, P(C.c_uint32),
                          C.c_double, C.c_double, C.c_uint64, P(Stats)]
    L.orc_run_batch.argtypes = [P(Network), C.c_int32, P(Topology), P(Params), P(C.c_uint32),
                                P(P(C.c_uint32)), C.c_double, C.c_uint64, P(Stats)]
    L.orc_ssa_run.argtypes = [P(Network), P(Topology), C.c_uint64, C.c_uint32, P(C.c_uint32),
                              C.c_double, P(C.c_uint64)]
    L.orc_ssa_batch.argtypes = [P(Network), P(Topology), C.c_uint64, C.c_int32, P(C.c_uint32),
                                C.c_double]
    L.orc_num_threads.restype = C.c_int
    if hasattr(L, "orc_mlp_forward"):
        _declare_mlp(L)


def _declare_mlp(L):
    P = C.POINTER
    L.orc_mlp_param_count.restype = C.c_int
    L.orc_mlp_forward.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float), C.c_int32, C.c_int32,
                                  C.c_int32, P(C.c_float)]
    L.orc_mlp_forward.restype = None
    L.orc_mlp_loss_grad.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float), C.c_int32, C.c_int32,
                                    C.c_int32, P(C.c_float), P(C.c_float), P(C.c_double)]
    L.orc_gelu.argtypes = [C.c_float]
    L.orc_gelu.restype = C.c_float
    L.orc_mlp_loss_grad.restype = C.c_double
    L.orc_adam_step.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float), C.c_int32,
                                C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float]
    L.orc_adam_step.restype = None
    L.orc_features.argtypes = [P(Network), P(Topology), P(Params), P(C.c_uint32), C.c_float,
                               P(C.c_float)]
    L.orc_violations.argtypes = [P(Network), P(Topology), P(Params), P(C.c_uint32),
                                 P(C.c_uint32), C.c_double, P(C.c_uint8)]
    L.orc_records.argtypes = [P(Network), P(Topology), P(Params), P(C.c_uint32), C.c_uint32,
                              C.c_uint32, C.c_float, C.c_double, P(C.c_float), P(C.c_float),
                              P(C.c_float)]


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Model:
    """Keeps numpy arrays alive and exposes the ctypes structs."""

    def __init__(self, net, topo):
        # net: dict(S,R,C,reac[S*R],prod[S*R],k[C*R],d[C*S]); topo: dict(N,row_ptr,col,w,comp)
        self.reac = np.ascontiguousarray(net["reac"], dtype=np.int32)
        self.prod = np.ascontiguousarray(net["prod"], dtype=np.int32)
        self.k = np.ascontiguousarray(net["k"], dtype=np.float64)
        self.d = np.ascontiguousarray(net["d"], dtype=np.float64)
        self.net = Network(net["S"], net["R"], net["C"], _ptr(self.reac, C.c_int32),
                           _ptr(self.prod, C.c_int32), _ptr(self.k, C.c_double),
                           _ptr(self.d, C.c_double))
        self.S, self.R = net["S"], net["R"]
        self.topo, self._keep = make_topology(topo)
        self.N = int(topo["N"])


def make_topology(topo):
    rp = np.ascontiguousarray(topo["row_ptr"], dtype=np.uint32)
    col = np.ascontiguousarray(topo["col"], dtype=np.uint32)
    keep = [rp, col]
    w = topo.get("w")
    comp = topo.get("comp")
    wp = C.POINTER(C.c_float)()
    cp = C.POINTER(C.c_int32)()
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        keep.append(w)
        wp = _ptr(w, C.c_float)
    if comp is not None:
        comp = np.ascontiguousarray(comp, dtype=np.int32)
        keep.append(comp)
        cp = _ptr(comp, C.c_int32)
    t = Topology(int(topo["N"]), int(len(col)), _ptr(rp, C.c_uint32), _ptr(col, C.c_uint32), wp, cp)
    return t, keep


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return list(o)


def poisson_many(lam, seed, n):
    out = np.empty(n, dtype=np.int64)
    lib().orc_poisson_many(lam, seed, n, _ptr(out, C.c_int64))
    return out


def poisson_inversion(lam, u0):
    """The oracle's draw for mean lam < 10 from packed word u0 (the search
    does not use the PTRS stream, so the counter is irrelevant)."""
    z = (C.c_uint32 * 4)()
    k = (C.c_uint32 * 2)()
    return int(lib().orc_poisson(lam, u0, z, k, 0))


def squeeze_poisson(lam, u0):
    """The kernels' fast path for one inversion draw: k, or -1 if left to
    the double search."""
    return int(lib().orc_sq_poisson(lam, u0))


def squeeze_decide(l, u, band):
    return int(lib().orc_sq_decide_band(l, u, band))


def squeeze_hop_check(tau, d, x, w, u0):
    """The kernels' single-precision hop draw (fast: k or -1) next to the
    oracle's draw from the double mean (exact: k, or -1 for mean >= 10)."""
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    u0 = np.ascontiguousarray(u0, dtype=np.uint32)
    n = len(tau)
    fast = np.empty(n, dtype=np.int32)
    exact = np.empty(n, dtype=np.int64)
    lib().orc_sq_hop_check(_ptr(tau, C.c_double), _ptr(d, C.c_double), _ptr(x, C.c_uint32),
                           _ptr(w, C.c_float), _ptr(u0, C.c_uint32), n, _ptr(fast, C.c_int32),
                           _ptr(exact, C.c_int64))
    return fast, exact


def squeeze_maxerr(b0, b1, stride=1):
    return float(lib().orc_sq_maxerr(b0, b1, stride))


def tau_all(model, x, eps, nc, seed, realization=0, step=0):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    N = model.N
    tp = np.empty(N)
    a0c = np.empty(N)
    t2 = np.empty(N)
    p = Params(eps, nc, seed)
    rc = lib().orc_tau_all(C.byref(model.net), C.byref(model.topo), C.byref(p),
                           _ptr(x, C.c_uint32), realization, step, _ptr(tp, C.c_double),
                           _ptr(a0c, C.c_double), _ptr(t2, C.c_double))
    assert rc == 0
    return tp, a0c, t2


def critical_choice(model, x, eps, nc, seed, v, steps, realization=0):
    """The oracle's critical-event choice at subvolume v for every step index
    in steps: arrays (kind, idx, dst)."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    steps = np.ascontiguousarray(steps, dtype=np.uint32)
    n = steps.size
    kind = np.empty(n, np.int32)
    idx = np.empty(n, np.int32)
    dst = np.empty(n, np.uint32)
    p = Params(eps, nc, seed)
    rc = lib().orc_critical_choice(C.byref(model.net), C.byref(model.topo), C.byref(p),
                                   _ptr(x, C.c_uint32), realization, v, n, _ptr(steps, C.c_uint32),
                                   _ptr(kind, C.c_int32), _ptr(idx, C.c_int32), _ptr(dst, C.c_uint32))
    assert rc == 0
    return kind, idx, dst


def run(model, x, eps, nc, seed, t_end, step_limit=2**63, realization=0, t_pause=float("inf"),
        stats=None):
    """Advance x (SoA uint32 [S*N]) in place; returns Stats."""
    assert x.dtype == np.uint32 and x.flags.c_contiguous
    st = stats if stats is not None else Stats()
    p = Params(eps, nc, seed)
    rc = lib().orc_run(C.byref(model.net), C.byref(model.topo), C.byref(p), realization,
                       _ptr(x, C.c_uint32), t_end, t_pause, step_limit, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle run failed rc={rc}")
    return st


def run_batch(net_model, topos, xs, eps, nc, seed, t_end, step_limit=2**63, ids=None):
    """Many realizations; topos is a list of topology dicts, xs a list of uint32 arrays."""
    n = len(xs)
    tarr = (Topology * n)()
    keep = []
    for i, t in enumerate(topos):
        tt, kk = make_topology(t)
        tarr[i] = tt
        keep.append(kk)
    xp = (C.POINTER(C.c_uint32) * n)(*[_ptr(x, C.c_uint32) for x in xs])
    st = (Stats * n)()
    idarr = (C.c_uint32 * n)(*(ids if ids is not None else range(n)))
    p = Params(eps, nc, seed)
    rc = lib().orc_run_batch(C.byref(net_model.net), n, tarr, C.byref(p), idarr, xp, t_end,
                             step_limit, st)
    if rc != 0:
        raise RuntimeError(f"oracle batch failed rc={rc}")
    return list(st)


def ssa_batch(model, x0, seed, nreal, t_end):
    S, N = model.S, model.N
    xs = np.tile(np.ascontiguousarray(x0, dtype=np.uint32), nreal).copy()
    rc = lib().orc_ssa_batch(C.byref(model.net), C.byref(model.topo), seed, nreal,
                             _ptr(xs, C.c_uint32), t_end)
    assert rc == 0
    return xs.reshape(nreal, S, N)


def num_threads():
    return lib().orc_num_threads()


# ---- learned tau reference ----
def mlp_param_count():
    return lib().orc_mlp_param_count()


def features(model, x, eps, nc, prev_logtau=0.0):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty((model.N, 12), dtype=np.float32)
    p = Params(eps, nc, 0)
    rc = lib().orc_features(C.byref(model.net), C.byref(model.topo), C.byref(p), _ptr(x, C.c_uint32),
                            prev_logtau, _ptr(out, C.c_float))
    assert rc == 0
    return out


def violations(model, x, xn, eps, nc, theta):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    xn = np.ascontiguousarray(xn, dtype=np.uint32)
    flags = np.empty(model.N, dtype=np.uint8)
    p = Params(eps, nc, 0)
    n = lib().orc_violations(C.byref(model.net), C.byref(model.topo), C.byref(p), _ptr(x, C.c_uint32),
                             _ptr(xn, C.c_uint32), theta, _ptr(flags, C.c_uint8))
    return n, flags


def records(model, x, eps, nc, seed, theta=1.0, realization=0, step=0, prev_logtau=0.0):
    """Training records of state x: features [N][12], log Cao tau' [N] and
    the trial-leap reject label [N]."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    N = model.N
    f = np.empty((N, 12), np.float32)
    y = np.empty(N, np.float32)
    lab = np.empty(N, np.float32)
    p = Params(eps, nc, seed)
    rc = lib().orc_records(C.byref(model.net), C.byref(model.topo), C.byref(p), _ptr(x, C.c_uint32),
                           realization, step, prev_logtau, theta, _ptr(f, C.c_float), _ptr(y, C.c_float),
                           _ptr(lab, C.c_float))
    assert rc == 0
    return f, y, lab


def gelu(z):
    """The networks' GELU (single precision), elementwise."""
    return np.array([lib().orc_gelu(float(v)) for v in np.ravel(z)], np.float32).reshape(np.shape(z))


def mlp_forward(w, norm, feats, bf16=True, act=1):
    w = np.ascontiguousarray(w, dtype=np.float32)
    norm = np.ascontiguousarray(norm, dtype=np.float32)
    feats = np.ascontiguousarray(feats, dtype=np.float32)
    out = np.empty((feats.shape[0], 2), dtype=np.float32)
    lib().orc_mlp_forward(_ptr(w, C.c_float), _ptr(norm, C.c_float), _ptr(feats, C.c_float),
                          feats.shape[0], 1 if bf16 else 0, act, _ptr(out, C.c_float))
    return out


def mlp_loss_grad(w, norm, feats, ylog, ycrit, bf16=True, act=1):
    w = np.ascontiguousarray(w, dtype=np.float32)
    norm = np.ascontiguousarray(norm, dtype=np.float32)
    feats = np.ascontiguousarray(feats, dtype=np.float32)
    ylog = np.ascontiguousarray(ylog, dtype=np.float32)
    ycrit = np.ascontiguousarray(ycrit, dtype=np.float32)
    g = np.empty(mlp_param_count(), dtype=np.float64)
    loss = lib().orc_mlp_loss_grad(_ptr(w, C.c_float), _ptr(norm, C.c_float), _ptr(feats, C.c_float),
                                   feats.shape[0], 1 if bf16 else 0, act, _ptr(ylog, C.c_float),
                                   _ptr(ycrit, C.c_float), _ptr(g, C.c_double))
    return loss, g


def adam_step(w, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    for a in (w, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    g = np.ascontiguousarray(g, dtype=np.float32)
    lib().orc_adam_step(_ptr(w, C.c_float), _ptr(m, C.c_float), _ptr(v, C.c_float), _ptr(g, C.c_float),
                        w.size, t, lr, b1, b2, eps)
