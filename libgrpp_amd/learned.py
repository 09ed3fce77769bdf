"""Learned radial-grid levels: the level predictor (an MLP trained and evaluated by the HIP kernels
of csrc/mlp.hip) and its use in a batched plan.

The unit of prediction is a row of a radial table: one (centre C, channel L, shell A, lambda, n_A)
combination.  A row's label is the statistics [max, mean, fraction >= 7] of the stopping levels of
its entries, recorded by the plain adaptive quadrature itself (the reference's nested-doubling rule,
radial_type2_integral.c:242-318; level k integrates 32 * 2^k - 1 Log3 points).  In a plan, the
predicted maximum orders the rows of each table so that 64 x 64 tiles gather entries of like
level; every entry still applies the reference's rule at every level of its tile (the embedded
check), and an entry not settled by the level its tile reaches continues adaptively (the
fallback).  Results therefore do not depend on the predictions, only the work does.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Tuple

import numpy as np

from . import configs
from . import lib as G
from .io import System

NIN, NOUT = 12, 3
_fp = C.POINTER(C.c_float)


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _bind():
    L = G.lib()
    if getattr(L, "_mlp_bound", False):
        return L
    L.libgrpp_mlp_create.argtypes = [C.c_uint, C.POINTER(C.c_void_p)]
    L.libgrpp_mlp_set_normalisation.argtypes = [C.c_void_p, _fp, _fp]
    L.libgrpp_mlp_params.argtypes = [C.c_void_p, _fp, _fp]
    L.libgrpp_mlp_gradient.argtypes = [C.c_void_p, C.c_int, _fp, _fp, _fp, _fp]
    L.libgrpp_mlp_train.argtypes = [C.c_void_p, C.c_int, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_uint, _fp]
    L.libgrpp_mlp_predict.argtypes = [C.c_void_p, C.c_int, _fp, _fp, C.c_int]
    L.libgrpp_mlp_destroy.argtypes = [C.c_void_p]
    L.libgrpp_mlp_destroy.restype = None
    L._mlp_bound = True
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class MLP:
    """12 -> 128 -> 128 -> 128 -> 3 SiLU perceptron on the GPU (libgrpp_mlp_*)."""

    def __init__(self, seed: int = 0):
        self.L = _bind()
        self.h = C.c_void_p()
        G._check(self.L.libgrpp_mlp_create(seed, C.byref(self.h)), "libgrpp_mlp_create")
        self.nparams = self.L.libgrpp_mlp_nparams()

    def set_normalisation(self, mean: np.ndarray, scale: np.ndarray):
        m, s = _f32(mean), _f32(scale)
        G._check(self.L.libgrpp_mlp_set_normalisation(self.h, _ptr(m), _ptr(s)), "libgrpp_mlp_set_normalisation")

    def params(self, set_to: np.ndarray = None) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        src = None if set_to is None else _f32(set_to)
        G._check(self.L.libgrpp_mlp_params(self.h, _ptr(out), None if src is None else _ptr(src)), "libgrpp_mlp_params")
        return out

    def gradient(self, X: np.ndarray, T: np.ndarray) -> Tuple[np.ndarray, float]:
        X, T = _f32(X), _f32(T)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float()
        G._check(self.L.libgrpp_mlp_gradient(self.h, X.shape[0], _ptr(X), _ptr(T), _ptr(g), C.byref(loss)),
                 "libgrpp_mlp_gradient")
        return g, loss.value

    def train(self, X: np.ndarray, T: np.ndarray, epochs: int = 40, batch: int = 256, lr: float = 2e-3,
              seed: int = 0) -> float:
        X, T = _f32(X), _f32(T)
        loss = C.c_float()
        G._check(self.L.libgrpp_mlp_train(self.h, X.shape[0], _ptr(X), _ptr(T), epochs, batch, lr, seed,
                                          C.byref(loss)), "libgrpp_mlp_train")
        return loss.value

    def predict(self, X: np.ndarray, bf16: bool = True) -> np.ndarray:
        X = _f32(X)
        Y = np.zeros((X.shape[0], NOUT), dtype=np.float32)
        G._check(self.L.libgrpp_mlp_predict(self.h, X.shape[0], _ptr(X), _ptr(Y), int(bf16)), "libgrpp_mlp_predict")
        return Y

    def close(self):
        if self.h:
            self.L.libgrpp_mlp_destroy(self.h)
            self.h = C.c_void_p()


def make_plan(workload, flags: int = 0):
    """a plan for a System, or for a list of Systems (independent molecules in one plan)"""
    if isinstance(workload, (list, tuple)):
        return G.MoleculesPlan(workload, flags)
    return G.Plan(workload, flags)


def row_labels(sysm) -> Tuple[np.ndarray, np.ndarray, dict]:
    """Features and stopping-level labels of every radial row of ``sysm`` (a System or a list of
    molecules), from one execution of the plain adaptive method."""
    plan = make_plan(sysm, G.BATCH_RECORD_LEVELS)
    try:
        plan.execute()
        return plan.features(), plan.row_levels(), plan.stats()
    finally:
        plan.close()


def training_set(systems: Iterable) -> Tuple[np.ndarray, np.ndarray]:
    xs, ts = [], []
    for s in systems:
        x, t, _ = row_labels(s)
        xs.append(x)
        ts.append(t)
    return np.concatenate(xs), np.concatenate(ts)


def train_predictor(X: np.ndarray, T: np.ndarray, epochs: int = 40, batch: int = 256, lr: float = 2e-3,
                    seed: int = 0) -> Tuple[MLP, float]:
    mlp = MLP(seed)
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    mlp.set_normalisation(mean, 1.0 / np.where(std > 1e-6, std, 1.0))
    loss = mlp.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, loss


def default_training_systems(n: int = 4, workload: str = "heavy_cluster"):
    """training systems of the workload's class with seeds 1..n (the measured workload uses seed 0);
    for ragged_molecules each system is a batch of 20 molecules"""
    if workload == "ragged_molecules":
        return [configs.ragged_molecules(seed=s, nmol=20) for s in range(1, n + 1)]
    return [configs.heavy_cluster(seed=s) for s in range(1, n + 1)]


def apply(plan: "G.Plan", mlp: MLP, bf16: bool = True) -> np.ndarray:
    """Predicts the level statistics of every row of ``plan`` and orders its radial tables by them."""
    pred = mlp.predict(plan.features(), bf16=bf16)
    plan.set_predictions(pred)
    return pred
