"""ctypes binding of libplis_hip.so (the C ABI in include/plis.h).

The GPU path is the only path: if the library is missing this module raises
on first use instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplis_hip.so")

PLIS_OK = 0
ERRORS = {-1: "PLIS_ERR_ARG", -2: "PLIS_ERR_HIP", -3: "PLIS_ERR_LEAPBACK", -4: "PLIS_ERR_NODEV",
          -5: "PLIS_ERR_STATE"}

# every entry point declared in include/plis.h
EXPORTS = ("plis_create", "plis_run", "plis_run_until", "plis_arm", "plis_iterate", "plis_profile",
           "plis_profile_kernels",
           "plis_set_state", "plis_get_state", "plis_get_stats", "plis_sync", "plis_debug_tau",
           "plis_launch_shape", "plis_redo_tiles", "plis_checkpoint_size", "plis_checkpoint", "plis_restore", "plis_autotune", "plis_destroy", "plis_last_error",
           "plis_device_count", "plis_mlp_set", "plis_mlp_get", "plis_mlp_forward",
           "plis_mlp_loss_grad", "plis_mlp_train", "plis_records", "plis_set_tau_mode",
           "plis_mlp_activation", "plis_mlp_forward_noalloc", "plis_debug_poison")
MLP_PARAMS = 5378


class Network(C.Structure):
    _fields_ = [("n_species", C.c_int32), ("n_reactions", C.c_int32), ("n_compartments", C.c_int32),
                ("reactants", C.POINTER(C.c_int32)), ("products", C.POINTER(C.c_int32)),
                ("k", C.POINTER(C.c_double)), ("d", C.POINTER(C.c_double))]


class Topology(C.Structure):
    _fields_ = [("n_subvolumes", C.c_uint32), ("n_edges", C.c_uint32),
                ("row_ptr", C.POINTER(C.c_uint32)), ("col", C.POINTER(C.c_uint32)),
                ("weight", C.POINTER(C.c_float)), ("compartment", C.POINTER(C.c_int32))]


class Params(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("n_critical", C.c_int32), ("waves_per_block", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("use_graph", C.c_int32), ("seed", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("t", C.c_double), ("tau", C.c_double), ("steps", C.c_uint64),
                ("leapbacks", C.c_uint64), ("firings", C.c_uint64), ("subvolume_leaps", C.c_uint64),
                ("crit_events", C.c_uint64), ("status", C.c_int32), ("pad", C.c_int32),
                ("learned_steps", C.c_uint64), ("learned_rejects", C.c_uint64),
                ("violations", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first "
                               "(there is no non-GPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = C.POINTER
    vp = C.c_void_p
    L.plis_create.argtypes = [P(Network), C.c_int32, P(Topology), P(P(C.c_uint32)), P(C.c_uint32),
                              P(Params), P(vp)]
    L.plis_run.argtypes = [vp, C.c_double, C.c_uint64]
    L.plis_run_until.argtypes = [vp, C.c_double, C.c_double, C.c_uint64]
    L.plis_arm.argtypes = [vp, C.c_double, C.c_double, C.c_uint64]
    L.plis_iterate.argtypes = [vp, C.c_uint32]
    L.plis_profile.argtypes = [vp, C.c_uint32, P(C.c_double)]
    L.plis_profile_kernels.argtypes = [vp, C.c_uint32, P(C.c_double)]
    L.plis_set_state.argtypes = [vp, C.c_int32, P(C.c_uint32)]
    L.plis_get_state.argtypes = [vp, C.c_int32, P(C.c_uint32)]
    L.plis_get_stats.argtypes = [vp, C.c_int32, P(Stats)]
    L.plis_sync.argtypes = [vp]
    L.plis_debug_tau.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.plis_launch_shape.argtypes = [vp, P(C.c_int32)]
    L.plis_redo_tiles.argtypes = [vp, P(C.c_uint64)]
    L.plis_checkpoint_size.argtypes = [vp, P(C.c_uint64)]
    L.plis_checkpoint.argtypes = [vp, C.c_void_p, C.c_uint64]
    L.plis_restore.argtypes = [vp, C.c_void_p, C.c_uint64]
    L.plis_autotune.argtypes = [vp, C.c_uint32]
    L.plis_destroy.argtypes = [vp]
    L.plis_destroy.restype = None
    L.plis_last_error.argtypes = [vp]
    L.plis_last_error.restype = C.c_char_p
    L.plis_device_count.argtypes = [C.c_char_p, C.c_int32]
    fp = P(C.c_float)
    L.plis_mlp_set.argtypes = [vp, fp, fp]
    L.plis_mlp_get.argtypes = [vp, fp]
    L.plis_mlp_forward.argtypes = [vp, fp, C.c_uint32, fp]
    L.plis_mlp_forward_noalloc.argtypes = [vp, fp, C.c_uint32, fp]
    L.plis_debug_poison.argtypes = [vp]
    L.plis_mlp_loss_grad.argtypes = [vp, fp, fp, fp, C.c_uint32, fp, P(C.c_double)]
    L.plis_mlp_train.argtypes = [vp, fp, fp, fp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, fp]
    L.plis_records.argtypes = [vp, C.c_int32, fp, fp, fp]
    L.plis_set_tau_mode.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_double]
    L.plis_mlp_activation.argtypes = [vp, C.c_int32]


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count():
    buf = C.create_string_buffer(256)
    n = lib().plis_device_count(buf, 256)
    return n, buf.value.decode(errors="replace")


class PlisError(RuntimeError):
    pass


class Context:
    """A batch of realizations of one reaction network on the GPU.

    net: dict(S, R, C, reac[S*R], prod[S*R], k[C*R], d[C*S])
    topos: list of dict(N, row_ptr, col, w=None, comp=None)
    x0: list of uint32 arrays (species-major [S*N])
    """

    def __init__(self, net, topos, x0, eps=0.03, nc=10, seed=0, ids=None, waves_per_block=0,
                 blocks_per_cu=0, use_graph=True):
        L = lib()
        self.S, self.R = int(net["S"]), int(net["R"])
        self._keep = []
        reac = self._k(np.asarray(net["reac"], dtype=np.int32))
        prod = self._k(np.asarray(net["prod"], dtype=np.int32))
        k = self._k(np.asarray(net["k"], dtype=np.float64))
        d = self._k(np.asarray(net["d"], dtype=np.float64))
        self.net = Network(self.S, self.R, int(net["C"]), _ptr(reac, C.c_int32),
                           _ptr(prod, C.c_int32), _ptr(k, C.c_double), _ptr(d, C.c_double))
        n = len(topos)
        self.n = n
        tarr = (Topology * n)()
        self.N = []
        for i, t in enumerate(topos):
            rp = self._k(np.asarray(t["row_ptr"], dtype=np.uint32))
            col = self._k(np.asarray(t["col"], dtype=np.uint32))
            wp = C.POINTER(C.c_float)()
            cp = C.POINTER(C.c_int32)()
            if t.get("w") is not None:
                wp = _ptr(self._k(np.asarray(t["w"], dtype=np.float32)), C.c_float)
            if t.get("comp") is not None:
                cp = _ptr(self._k(np.asarray(t["comp"], dtype=np.int32)), C.c_int32)
            tarr[i] = Topology(int(t["N"]), int(len(col)), _ptr(rp, C.c_uint32),
                               _ptr(col, C.c_uint32), wp, cp)
            self.N.append(int(t["N"]))
        xs = [self._k(np.asarray(x, dtype=np.uint32)) for x in x0]
        for x, N in zip(xs, self.N):
            if x.size != self.S * N:
                raise ValueError("x0 size must be S*N")
        xp = (C.POINTER(C.c_uint32) * n)(*[_ptr(x, C.c_uint32) for x in xs])
        idp = None
        if ids is not None:
            idp = (C.c_uint32 * n)(*ids)
        p = Params(eps, nc, waves_per_block, blocks_per_cu, 1 if use_graph else 0, seed)
        h = C.c_void_p()
        rc = L.plis_create(C.byref(self.net), n, tarr, xp, idp, C.byref(p), C.byref(h))
        if rc != PLIS_OK:
            raise PlisError(f"plis_create: {ERRORS.get(rc, rc)}: {L.plis_last_error(None).decode()}")
        self.h = h
        self._keep = []  # host arrays were copied

    def _k(self, a):
        a = np.ascontiguousarray(a)
        self._keep.append(a)
        return a

    def _chk(self, rc, what):
        if rc != PLIS_OK:
            raise PlisError(f"{what}: {ERRORS.get(rc, rc)}: {lib().plis_last_error(self.h).decode()}")

    def run(self, t_end, step_limit=2**63, t_pause=float("inf")):
        self._chk(lib().plis_run_until(self.h, t_end, t_pause, step_limit), "plis_run_until")

    def arm(self, t_end=float("inf"), step_limit=2**63, t_pause=float("inf")):
        self._chk(lib().plis_arm(self.h, t_end, t_pause, step_limit), "plis_arm")

    def iterate(self, n):
        self._chk(lib().plis_iterate(self.h, n), "plis_iterate")

    def profile(self, n):
        """[total, tau, decide, leap, commit] milliseconds summed over n iterations."""
        out = (C.c_double * 5)()
        self._chk(lib().plis_profile(self.h, n, out), "plis_profile")
        return list(out)

    def profile_kernels(self, n):
        """profile() plus the leap pass by kernel: [..., phase_b, redo, finish]."""
        out = (C.c_double * 8)()
        self._chk(lib().plis_profile_kernels(self.h, n, out), "plis_profile_kernels")
        return list(out)

    def sync(self):
        self._chk(lib().plis_sync(self.h), "plis_sync")

    def autotune(self, iters=3):
        self._chk(lib().plis_autotune(self.h, iters), "plis_autotune")

    def launch_shape(self):
        out = (C.c_int32 * 3)()
        self._chk(lib().plis_launch_shape(self.h, out), "plis_launch_shape")
        return {"waves_per_block": out[0], "blocks": out[1], "tiles": out[2]}

    def checkpoint(self):
        """Every realization's counts, clock, indices and statistics as bytes."""
        n = C.c_uint64()
        self._chk(lib().plis_checkpoint_size(self.h, C.byref(n)), "plis_checkpoint_size")
        buf = C.create_string_buffer(n.value)
        self._chk(lib().plis_checkpoint(self.h, buf, n.value), "plis_checkpoint")
        return buf.raw

    def restore(self, blob):
        """Continue from a checkpoint of a context with the same model and seed."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._chk(lib().plis_restore(self.h, buf, len(blob)), "plis_restore")

    def debug_poison(self):
        """Sentinels in every buffer a step writes before it reads (initcheck stand-in)."""
        self._chk(lib().plis_debug_poison(self.h), "plis_debug_poison")

    def redo_tiles(self):
        """Tiles the edge-parallel leap passes have left to their redo pass
        since the context was made."""
        n = C.c_uint64()
        self._chk(lib().plis_redo_tiles(self.h, C.byref(n)), "plis_redo_tiles")
        return n.value

    def get_state(self, r=0):
        x = np.empty(self.S * self.N[r], dtype=np.uint32)
        self._chk(lib().plis_get_state(self.h, r, _ptr(x, C.c_uint32)), "plis_get_state")
        return x

    def set_state(self, r, x):
        x = np.ascontiguousarray(x, dtype=np.uint32)
        self._chk(lib().plis_set_state(self.h, r, _ptr(x, C.c_uint32)), "plis_set_state")

    def stats(self, r=0):
        st = Stats()
        self._chk(lib().plis_get_stats(self.h, r, C.byref(st)), "plis_get_stats")
        return st.as_dict()

    def debug_tau(self):
        n = sum(self.N)
        tp, a0c, t2 = np.empty(n), np.empty(n), np.empty(n)
        self._chk(lib().plis_debug_tau(self.h, _ptr(tp, C.c_double), _ptr(a0c, C.c_double),
                                       _ptr(t2, C.c_double)), "plis_debug_tau")
        return tp, a0c, t2

    # ---- learned tau ----
    def mlp_set(self, w32, norm):
        w = np.ascontiguousarray(w32, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        assert w.size == MLP_PARAMS and nm.size == 26
        self._chk(lib().plis_mlp_set(self.h, _ptr(w, C.c_float), _ptr(nm, C.c_float)), "plis_mlp_set")

    def mlp_get(self):
        w = np.empty(MLP_PARAMS, dtype=np.float32)
        self._chk(lib().plis_mlp_get(self.h, _ptr(w, C.c_float)), "plis_mlp_get")
        return w

    def mlp_forward(self, feats):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.empty((f.shape[0], 2), dtype=np.float32)
        self._chk(lib().plis_mlp_forward(self.h, _ptr(f, C.c_float), f.shape[0], _ptr(out, C.c_float)),
                  "plis_mlp_forward")
        return out

    def mlp_forward_noalloc(self, feats):
        """plis_mlp_forward through the library's fixed scratch (no device
        allocation; at most 16384 rows)."""
        f = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.empty((f.shape[0], 2), dtype=np.float32)
        self._chk(lib().plis_mlp_forward_noalloc(self.h, _ptr(f, C.c_float), f.shape[0], _ptr(out, C.c_float)),
                  "plis_mlp_forward_noalloc")
        return out

    def mlp_loss_grad(self, feats, ylog, ycrit):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(ylog, dtype=np.float32)
        c = np.ascontiguousarray(ycrit, dtype=np.float32)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_double()
        self._chk(lib().plis_mlp_loss_grad(self.h, _ptr(f, C.c_float), _ptr(y, C.c_float), _ptr(c, C.c_float),
                                           f.shape[0], _ptr(g, C.c_float), C.byref(loss)), "plis_mlp_loss_grad")
        return loss.value, g

    def mlp_train(self, feats, ylog, ycrit, steps, batch, lr, seed=1):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(ylog, dtype=np.float32)
        c = np.ascontiguousarray(ycrit, dtype=np.float32)
        losses = np.empty(steps, dtype=np.float32)
        self._chk(lib().plis_mlp_train(self.h, _ptr(f, C.c_float), _ptr(y, C.c_float), _ptr(c, C.c_float),
                                       f.shape[0], steps, batch, lr, seed, _ptr(losses, C.c_float)),
                  "plis_mlp_train")
        return losses

    def records(self, r=0):
        N = self.N[r]
        f = np.empty((N, 12), dtype=np.float32)
        y = np.empty(N, dtype=np.float32)
        c = np.empty(N, dtype=np.float32)
        self._chk(lib().plis_records(self.h, r, _ptr(f, C.c_float), _ptr(y, C.c_float), _ptr(c, C.c_float)),
                  "plis_records")
        return f, y, c

    def mlp_activation(self, act):
        """0 = ReLU, 1 = GELU (default)."""
        self._chk(lib().plis_mlp_activation(self.h, int(act)), "plis_mlp_activation")

    def set_tau_mode(self, learned, kappa=1.0, theta=1.0, viol_rate_max=1.0):
        self._chk(lib().plis_set_tau_mode(self.h, 1 if learned else 0, kappa, theta, viol_rate_max),
                  "plis_set_tau_mode")

    def close(self):
        if getattr(self, "h", None):
            lib().plis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
