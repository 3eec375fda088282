"""Learned stepper handle (message-passing correction of one large RK step
that stands in for 4 RK4 steps).

Thin wrapper over the mfree_surrogate_* / mfree_mpnn_* entry points of
include/mfree_iwf_amd.h; training, inference and the judged rollout all run
in the gfx950 library.
"""
import ctypes as C

import numpy as np

from . import _capi

NF = 16       # node features
NOUT = 8      # outputs: dx, dy, dvx, dvy, dSxx, dSxy, dSyy, drho (dimensionless)
MAX_DEG = 64  # edge slots per node


def _fp(a):
    return a.ctypes.data_as(_capi.fptr)


def _u32(a):
    return a.ctypes.data_as(_capi.u32ptr)


def param_count():
    return int(_capi.lib().mfree_mpnn_param_count())


def layout():
    """[(rows, cols), ...] of the flat parameter vector (row-major, out x in)."""
    L = _capi.lib()
    nt = L.mfree_mpnn_layout(None)
    sh = np.zeros(2 * nt, dtype=np.int32)
    L.mfree_mpnn_layout(sh.ctypes.data_as(_capi.i32ptr))
    return [(int(sh[2 * k]), int(sh[2 * k + 1])) for k in range(nt)]


def forward_backward(row_ptr, col, pos3, x, t, w, device=0, train=True):
    """One forward + backward on the GPU over nodes at pos3 = (x, y, h):
    returns (y [n x 8], loss, grad).  With train=False only the inference
    forward runs (t is ignored): returns (y, None, None)."""
    n = len(row_ptr) - 1
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    pos3 = np.ascontiguousarray(pos3, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    y = np.zeros((n, NOUT), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    grad = np.zeros(param_count(), dtype=np.float32)
    if len(col) == 0:
        col = np.zeros(1, dtype=np.uint32)
    if not train:
        st = _capi.lib().mfree_mpnn_forward_backward(device, n, _u32(row_ptr), _u32(col), _fp(pos3), _fp(x), None,
                                                     _fp(w), _fp(y), None, None)
        _capi.check(st, what="mfree_mpnn_forward_backward")
        return y, None, None
    st = _capi.lib().mfree_mpnn_forward_backward(device, n, _u32(row_ptr), _u32(col), _fp(pos3), _fp(x), _fp(t),
                                                 _fp(w), _fp(y), _fp(loss), _fp(grad))
    _capi.check(st, what="mfree_mpnn_forward_backward")
    return y, float(loss[0]), grad


class Surrogate:
    def __init__(self, device=0, seed=1, max_batch_nodes=1 << 20, base_order=4):
        L = _capi.lib()
        o = _capi.SurrogateOptions()
        L.mfree_surrogate_default_options(C.byref(o))
        o.device = device
        o.seed = seed
        o.max_batch_nodes = max_batch_nodes
        o.base_order = base_order
        h = C.c_void_p()
        _capi.check(L.mfree_surrogate_create(C.byref(o), C.byref(h)), what="mfree_surrogate_create")
        self._h = h
        self._L = L

    def close(self):
        if self._h:
            self._L.mfree_surrogate_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, what):
        if st != _capi.OK:
            msg = self._L.mfree_surrogate_last_error(self._h).decode()
            raise _capi.MfreeError(st, f"{what}: {msg or self._L.mfree_status_string(st).decode()}")

    def collect(self, solver, ncapture, gap=0):
        self._chk(self._L.mfree_surrogate_collect(self._h, solver._h, ncapture, gap), "collect")

    @property
    def snapshots(self):
        return int(self._L.mfree_surrogate_snapshots(self._h))

    @property
    def dataset_nodes(self):
        return int(self._L.mfree_surrogate_dataset_nodes(self._h))

    def train(self, iters, lr=1e-3):
        loss = np.zeros(max(iters, 1), dtype=np.float32)
        self._chk(self._L.mfree_surrogate_train(self._h, iters, lr, _fp(loss)), "train")
        return loss[:iters]

    def record_velocity(self, solver):
        self._chk(self._L.mfree_surrogate_record_velocity(self._h, solver._h), "record_velocity")

    def step(self, solver, use_net=True):
        self._chk(self._L.mfree_surrogate_step(self._h, solver._h, 1 if use_net else 0), "step")

    def rollout(self, solver, nsteps, group=16, threshold=1e-4):
        js = _capi.JudgeStats()
        self._chk(self._L.mfree_surrogate_rollout(self._h, solver._h, nsteps, group, threshold, C.byref(js)),
                  "rollout")
        return {f: getattr(js, f) for f, _ in js._fields_}

    def params(self):
        w = np.zeros(param_count(), dtype=np.float32)
        norm = np.zeros(2 * NF + 2 * NOUT, dtype=np.float32)
        self._chk(self._L.mfree_surrogate_get_params(self._h, _fp(w), _fp(norm)), "get_params")
        return w, norm

    def set_params(self, w=None, norm=None):
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        na = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self._chk(self._L.mfree_surrogate_set_params(self._h, None if wa is None else _fp(wa),
                                                     None if na is None else _fp(na)), "set_params")

    def graph(self, solver, cap=None):
        """(row_ptr, col, edge3) of the learned step's graph by global index."""
        n = int(self._L.mfree_num_particles(solver._h))
        cap = cap if cap is not None else n * MAX_DEG
        row = np.zeros(n + 1, dtype=np.uint32)
        col = np.zeros(cap, dtype=np.uint32)
        e3 = np.zeros((cap, 3), dtype=np.float32)
        ne = C.c_uint64(0)
        self._chk(self._L.mfree_surrogate_graph(self._h, solver._h, _u32(row), _u32(col), _fp(e3), cap,
                                                C.byref(ne)), "graph")
        return row, col[:ne.value], e3[:ne.value]
