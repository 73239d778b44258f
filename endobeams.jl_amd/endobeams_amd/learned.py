"""Learned Newton warm start (include/endobeams_learned.h): the network, its training on the
solver's own trajectories, and its installation as the predictor's alternative start.

Workflow: run the solver on a few structures (`collect`), fit input/output scales, train the
network on the device (`Network.train`: the library's forward, back-propagation and Adam kernels),
then `WarmStart.install(engine)`. The reference predictor remains the fallback per instance, so
Newton converges to the same tolerance with or without the network.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

from . import abi
from .engine import EBError, load_hip_library

WIDTHS = (abi.LEARNED_INPUTS, 128, 128, 128, abi.LEARNED_OUTPUTS)   # three hidden layers of 128


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Network:
    """tanh MLP in FP32 on the device (eb_mlp_*)."""

    def __init__(self, widths: Sequence[int] = WIDTHS, seed: int = 0, lib: Optional[C.CDLL] = None):
        self.lib = lib or load_hip_library()
        self.fn = abi.bind(self.lib, "eb_")
        self.widths = tuple(int(w) for w in widths)
        w = np.asarray(self.widths, dtype=np.int32)
        h = C.c_void_p()
        st = self.fn["mlp_create"](w.ctypes.data_as(C.POINTER(C.c_int32)), len(w), seed, C.byref(h))
        if st != abi.EB_OK:
            raise EBError(f"eb_mlp_create failed ({st})")
        self.h = h
        n = C.c_int64(0)
        self._check(self.fn["mlp_num_params"](self.h, C.byref(n)), "num_params")
        self.n_params = n.value

    def _check(self, st, what):
        if st != abi.EB_OK:
            msg = self.fn["mlp_last_error"](self.h)
            raise EBError(f"eb_mlp_{what} failed ({st}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.fn["mlp_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._check(self.fn["mlp_get_params"](self.h, _fp(out), out.size), "get_params")
        return out

    @params.setter
    def params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.fn["mlp_set_params"](self.h, _fp(a), a.size), "set_params")

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], self.widths[-1]), np.float32)
        self._check(self.fn["mlp_forward"](self.h, _fp(X), X.shape[0], _fp(Y)), "forward")
        return Y

    def loss_gradient(self, X, T) -> Tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float(0)
        g = np.empty(self.n_params, np.float32)
        self._check(self.fn["mlp_loss_gradient"](self.h, _fp(X), _fp(T), X.shape[0], C.byref(loss), _fp(g)),
                    "loss_gradient")
        return loss.value, g

    def train(self, X, T, epochs=20, batch=4096, lr=1e-3, seed=0) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        self._check(self.fn["mlp_set_dataset"](self.h, _fp(X), _fp(T), X.shape[0]), "set_dataset")
        losses = np.zeros(epochs, np.float32)
        self._check(self.fn["mlp_train"](self.h, epochs, batch, lr, seed, _fp(losses)), "train")
        return losses


def sample(engine) -> Optional[Tuple[np.ndarray, np.ndarray]]:
    """(inputs, target corrections) of the step just converged, unscaled; None before three
    converged steps are recorded."""
    n = engine.sizes.n_nodes
    X = np.empty((n, abi.LEARNED_INPUTS), np.float32)
    T = np.empty((n, abi.LEARNED_OUTPUTS), np.float32)
    st = engine.fn["learned_sample"](engine.h, _fp(X), _fp(T))
    if st == abi.EB_ERR_STATE:
        return None
    engine._check(st, "learned_sample")
    return X, T


def collect(engine, params, steps: int) -> Tuple[np.ndarray, np.ndarray]:
    """Runs the reference time loop for `steps` converged steps and records a sample after each
    (from the fourth on)."""
    from .simulation import Stepper
    xs, ts = [], []

    def on_converged(step, t, eng):
        s = sample(eng)
        if s is not None:
            xs.append(s[0])
            ts.append(s[1])

    st = Stepper(engine, params)
    st.initialize()
    while st.stats.steps < steps and not st.finished:
        st.advance(on_converged)
    return np.concatenate(xs), np.concatenate(ts)


class WarmStart:
    """A trained network with the scales of its inputs and outputs."""

    def __init__(self, net: Network, in_scale, out_scale):
        self.net = net
        self.in_scale = np.ascontiguousarray(in_scale, dtype=np.float32)
        self.out_scale = np.ascontiguousarray(out_scale, dtype=np.float32)

    @staticmethod
    def fit_scales(X, T):
        """Inputs to unit spread per column group (the predictor's increment and the three recorded
        increments share one scale per component, contact force per component); outputs to unit
        spread per component."""
        X = np.asarray(X, np.float64)
        T = np.asarray(T, np.float64)
        s_in = np.ones(abi.LEARNED_INPUTS)
        for c in range(6):
            cols = [c, 6 + c, 12 + c, 18 + c]
            sd = np.sqrt(np.mean(X[:, cols] ** 2))
            s_in[cols] = 1.0 / sd if sd > 0 else 1.0
        for c in range(24, 27):
            sd = np.sqrt(np.mean(X[:, c] ** 2))
            s_in[c] = 1.0 / sd if sd > 0 else 1.0
        s_out = np.sqrt(np.mean(T ** 2, axis=0))
        s_out[s_out == 0] = 1.0
        return s_in.astype(np.float32), s_out.astype(np.float32)

    @classmethod
    def train(cls, X, T, epochs=20, batch=4096, lr=1e-3, seed=0, widths=WIDTHS):
        s_in, s_out = cls.fit_scales(X, T)
        net = Network(widths, seed=seed)
        losses = net.train(X * s_in, T / s_out, epochs=epochs, batch=batch, lr=lr, seed=seed)
        ws = cls(net, s_in, s_out)
        ws.losses = losses
        return ws

    def install(self, engine):
        engine._check(engine.fn["set_learned_predictor"](engine.h, self.net.h, _fp(self.in_scale),
                                                        _fp(self.out_scale)), "set_learned_predictor")


def uninstall(engine):
    engine._check(engine.fn["set_learned_predictor"](engine.h, None, None, None), "set_learned_predictor")


def set_backoff(engine, max_skip: int):
    """Longest run of predictions that skip a network whose start keeps losing (0: evaluate it
    at every prediction)."""
    engine._check(engine.fn["set_learned_backoff"](engine.h, int(max_skip)), "set_learned_backoff")


def learned_increment(engine) -> np.ndarray:
    """The increment ΔD of the network's start at the last prediction that evaluated it."""
    out = np.zeros(engine.sizes.ndofs)
    engine._check(engine.fn["get_learned_increment"](engine.h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size),
                  "get_learned_increment")
    return out


def predictor_stats(engine) -> np.ndarray:
    """[network start kept, reference start kept after evaluating both, network skipped]"""
    out = np.zeros(3, np.int64)
    engine._check(engine.fn["get_predictor_stats"](engine.h, out.ctypes.data_as(C.POINTER(C.c_int64))),
                  "get_predictor_stats")
    return out
