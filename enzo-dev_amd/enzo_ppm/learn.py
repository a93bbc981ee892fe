"""Learned two-shock starting pressure through the C ABI (hipPPMLearn*,
include/enzo_ppm.h).

The network is 6 -> 32 -> 32 -> 1, FP32, tanh hidden layers, predicting the
correction c = log(p* / p_ac) to the reference's acoustic start p_ac from
six dimensionless interface features; its inputs are standardised with the
training set's mean and 1/std, and the start is p_ac e^c only where the
acoustic wave strength passes the gate g (|p_ac - p_l| > g p_l or
|p_ac - p_r| > g p_r) -- all of which travel with the weights.  Training,
gradients, predictions and the Newton statistics run in the library's HIP
kernels; there is no CPU path.

Harvest records (DeviceGrid.get_harvest) are rows of 10 float64 values:
rho_l, rho_r, p_l, p_r, u_l, u_r, p*, starting pressure, Newton updates,
restarted.
"""
import ctypes as C

import numpy as np

from . import abi, device

N_IN, N_HIDDEN = 6, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HIDDEN * N_IN
OFF_W2 = OFF_B1 + N_HIDDEN
OFF_B2 = OFF_W2 + N_HIDDEN * N_HIDDEN
OFF_W3 = OFF_B2 + N_HIDDEN
OFF_B3 = OFF_W3 + N_HIDDEN
N_PARAMS = OFF_B3 + 1            # trained parameters
OFF_MEAN = N_PARAMS
OFF_INV_STD = OFF_MEAN + N_IN
OFF_GATE = OFF_INV_STD + N_IN    # |c| threshold for using the network
N_FLOATS = OFF_GATE + 1          # whole block
RECORD = 10


def _lib(lib):
    L = lib or abi.load_library()
    if L.hipPPMLearnParameterCount() != N_FLOATS:
        raise RuntimeError("library's network layout differs from enzo_ppm.learn")
    return L


def _check(L, st, what):
    if st != abi.SUCCESS:
        raise RuntimeError(f"{what} failed with status {st}: {L.hipPPMLearnLastError().decode()}")


def _records(records):
    r = np.ascontiguousarray(records, dtype=np.float64)
    if r.ndim != 2 or r.shape[1] != RECORD:
        raise ValueError("records must be an (n, 10) float64 array")
    return r


def _fptr(w):
    return w.ctypes.data_as(C.c_void_p)


def init_weights(seed=0):
    """Glorot-uniform weights, zero biases, identity standardisation, gate 0."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_FLOATS, np.float32)

    def glorot(n_out, n_in):
        a = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, size=n_out * n_in).astype(np.float32)

    w[OFF_W1:OFF_B1] = glorot(N_HIDDEN, N_IN)
    w[OFF_W2:OFF_B2] = glorot(N_HIDDEN, N_HIDDEN)
    w[OFF_W3:OFF_B3] = glorot(1, N_HIDDEN)
    w[OFF_INV_STD:OFF_GATE] = 1.0
    return w


def harvest(grid, glob, steps=1, keep=1, cap=1 << 22, cfl=0.4, lib=None, weights=None,
            boundary=True):
    """`steps` device-resident steps of `grid` (periodic ghost fill unless
    `boundary` is false -- then the ghosts stay as given --, Courant dt,
    rotated sweeps) with the harvest on; `grid` is updated in place.
    Returns (records, interfaces selected).  With `weights` the sweeps use
    the learned start and the records show its Newton updates."""
    dg = device.DeviceGrid(grid, glob, lib=lib)
    out = []
    seen = 0
    try:
        if weights is not None:
            dg.set_learned(weights)
        dg.set_harvest(cap, keep)
        for cycle in range(steps):
            if boundary:
                dg.periodic_boundary()
            dg.pressure_and_dt(cfl, want_host=False)
            dg.step(cycle)
            r, s = dg.get_harvest(cap)
            out.append(r)
            seen += s
        if dg.download() != abi.SUCCESS:
            raise RuntimeError("harvest run failed")
    finally:
        dg.close()
    return np.concatenate(out), seen


def train(records, gamma, epochs=4, batch=4096, lr=2e-3, seed=0, weights=None, lib=None):
    """Minibatch Adam on the device.  Returns (weights, mse after each epoch);
    the gate is left as given (0 from init_weights): see fit_gate."""
    L = _lib(lib)
    r = _records(records)
    w = np.array(init_weights(seed) if weights is None else weights, dtype=np.float32, copy=True)
    mse = np.zeros(max(1, epochs))
    _check(L, L.hipPPMLearnTrain(abi.dptr(r), len(r), gamma, epochs, batch, lr, seed, _fptr(w),
                                 abi.dptr(mse)), "LearnTrain")
    return w, mse[:epochs]


def gradient(records, gamma, weights, lib=None):
    """Full-batch gradient of the mean squared error (N_PARAMS values) and
    the mean squared error, from the training kernels."""
    L = _lib(lib)
    r = _records(records)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    g = np.zeros(N_PARAMS)
    m = C.c_double(0.0)
    _check(L, L.hipPPMLearnGradient(abi.dptr(r), len(r), gamma, _fptr(w), abi.dptr(g),
                                    C.byref(m)), "LearnGradient")
    return g, m.value


def predict(records, gamma, weights, lib=None):
    """(network outputs c = log(p*/p_ac) as float32, mean squared error)."""
    L = _lib(lib)
    r = _records(records)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    y = np.zeros(len(r), np.float32)
    m = C.c_double(0.0)
    _check(L, L.hipPPMLearnPredict(abi.dptr(r), len(r), gamma, _fptr(w), _fptr(y), C.byref(m)),
           "LearnPredict")
    return y, m.value


def newton_stats(records, gamma, pmin, weights, lib=None):
    """Two-shock solve of every record from both starts, on the device.
    Returns the per-record array (n, 7) and a summary dict."""
    L = _lib(lib)
    r = _records(records)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = np.zeros((len(r), 7))
    _check(L, L.hipPPMLearnNewtonStats(abi.dptr(r), len(r), gamma, pmin, _fptr(w),
                                       abi.dptr(out)), "LearnNewtonStats")
    s = summarize(out)
    s["network_evaluated_fraction"] = float(evaluated(r, out, w[OFF_GATE], pmin).mean())
    return out, s


def evaluated(records, out, gate, pmin=1e-20):
    """Which records pass the gate (the network is evaluated there): the
    floored acoustic start out[:, 6] against the left and right pressures."""
    p_ac, pl, pr = out[:, 6], records[:, 2], records[:, 3]
    return (p_ac > pmin) & ((np.abs(p_ac - pl) > gate * pl) | (np.abs(p_ac - pr) > gate * pr))


def fit_gate(records, gamma, pmin, weights,
             candidates=(0.0, 1e-3, 3e-3, 1e-2, 3e-2, 0.1, 0.3, 1.0), rel_tol=0.01, lib=None):
    """Gate choice on `records`: the largest candidate whose mean Newton
    updates are within `rel_tol` of the fewest any candidate gives (nearly
    the whole gain for the fewest network evaluations).  The closed gate
    (+inf: the network never evaluated, the reference start everywhere, the
    plain kernels) is a candidate too, so a network that cannot save
    `rel_tol` of the updates is switched off.  Returns (weights with that
    gate, {gate: (mean updates, fraction evaluated)})."""
    w = np.array(weights, dtype=np.float32, copy=True)
    r = np.ascontiguousarray(records, dtype=np.float64)
    table = {}
    for c in candidates:
        w[OFF_GATE] = c
        out, s = newton_stats(r, gamma, pmin, w, lib=lib)
        table[float(c)] = (s["newton_updates_learned_start"]["mean"],
                           float(evaluated(r, out, np.float32(c), pmin).mean()))
    table[float("inf")] = (s["newton_updates_reference_start"]["mean"], 0.0)
    best = min(v[0] for v in table.values())
    w[OFF_GATE] = max(c for c, v in table.items() if v[0] <= best * (1.0 + rel_tol))
    return w, table


def summarize(out):
    it_ref, it_learn, restarted = out[:, 0], out[:, 1], out[:, 2]
    p_ref, p_learn, start, start_ref = out[:, 3], out[:, 4], out[:, 5], out[:, 6]
    rel = np.abs(p_learn - p_ref) / np.abs(p_ref)
    start_err = np.abs(start - p_ref) / np.abs(p_ref)
    start_ref_err = np.abs(start_ref - p_ref) / np.abs(p_ref)
    return {
        "interfaces": int(len(out)),
        "newton_updates_reference_start": {"mean": float(it_ref.mean()),
                                           "p99": float(np.percentile(it_ref, 99))},
        "newton_updates_learned_start": {"mean": float(it_learn.mean()),
                                         "p99": float(np.percentile(it_learn, 99))},
        "restarted_fraction": float(restarted.mean()),
        "pstar_rel_error": {"max": float(rel.max()), "mean": float(rel.mean())},
        "learned_start_rel_error": {"median": float(np.median(start_err)),
                                    "p99": float(np.percentile(start_err, 99))},
        "reference_start_rel_error": {"median": float(np.median(start_ref_err)),
                                      "p99": float(np.percentile(start_ref_err, 99))},
    }
