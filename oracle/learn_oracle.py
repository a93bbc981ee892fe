"""ORACLE -- test infrastructure only (tests/, smoke check, bench CPU baseline).

Float64 numpy restatement of the learned two-shock start's network: the six
interface features, the target (the correction log(p* / p_ac) to the
reference's acoustic start p_ac, twoshock.F:156-170, clipped), the
standardised 6 -> 32 -> 32 -> 1 tanh forward pass, the mean squared error and
its gradient by hand-written backpropagation.  The library's FP32 HIP
kernels (enzo-dev_amd/csrc/ppm_learn.hip) are checked against it.

The reference has no counterpart (twoshock.F:141-254 always starts from the
acoustic estimate); what pins this restatement is the mathematics: the
backward pass is checked against central finite differences of the forward
pass (tests/test_learn_oracle.py), and the features and targets against
their closed forms on hand-computed states.

Weight block layout (float32 on the library side): W1[32][6] b1[32]
W2[32][32] b2[32] W3[32] b3, mean[6], inv_std[6], gate.
"""
import ctypes as C

import numpy as np

import oracle

N_IN, N_HIDDEN = 6, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HIDDEN * N_IN
OFF_W2 = OFF_B1 + N_HIDDEN
OFF_B2 = OFF_W2 + N_HIDDEN * N_HIDDEN
OFF_W3 = OFF_B2 + N_HIDDEN
OFF_B3 = OFF_W3 + N_HIDDEN
N_PARAMS = OFF_B3 + 1
OFF_MEAN = N_PARAMS
OFF_INV_STD = OFF_MEAN + N_IN
OFF_GATE = OFF_INV_STD + N_IN
N_FLOATS = OFF_GATE + 1
Y_CLIP = 8.0   # targets clipped to [-Y_CLIP, Y_CLIP]


def acoustic_start(records, gamma):
    """p_ac = (c_r p_l + c_l p_r + c_r c_l (u_l - u_r)) / (c_r + c_l) with the
    Lagrangian wave speeds c = sqrt(gamma p rho) (twoshock.F:156-170), before
    the floor."""
    r = np.asarray(records, dtype=np.float64)
    dl, dr, pl, pr, ul, ur = (r[:, k] for k in range(6))
    cl = np.sqrt(gamma * pl * dl)
    cr = np.sqrt(gamma * pr * dr)
    return (cr * pl + cl * pr + cr * cl * (ul - ur)) / (cr + cl)


def learnable(records, gamma):
    """Records training uses: positive states, p* and acoustic start."""
    r = np.asarray(records, dtype=np.float64)
    ok = np.all(r[:, [0, 1, 2, 3, 6]] > 0, axis=1)
    with np.errstate(invalid="ignore"):
        return ok & (acoustic_start(r, gamma) > 0)


def features(records, gamma, fp32=False):
    """X[n, 6]: log(p_r/p_l), log(rho_r/rho_l), (u_l-u_r)/c_l, (u_l-u_r)/c_r,
    c_l/(c_l+c_r), c_r/(c_l+c_r) with c = sqrt(gamma p / rho); targets
    y = log(p*/p_ac) clipped to [-Y_CLIP, Y_CLIP] (records: learnable ones).
    records: rows (rho_l, rho_r, p_l, p_r, u_l, u_r, p*, ...).
    fp32: round as the library's kernels do -- ratios and quotients in
    float64, rounded to float32, logarithms taken in float32."""
    r = np.asarray(records, dtype=np.float64)
    dl, dr, pl, pr, ul, ur, ps = (r[:, k] for k in range(7))
    cl = np.sqrt(gamma * pl / dl)
    cr = np.sqrt(gamma * pr / dr)
    du = ul - ur
    pac = acoustic_start(r, gamma)
    if fp32:
        f = np.float32
        X = np.stack([np.log((pr / pl).astype(f)), np.log((dr / dl).astype(f)),
                      (du / cl).astype(f), (du / cr).astype(f), (cl / (cl + cr)).astype(f),
                      (cr / (cl + cr)).astype(f)], axis=1).astype(np.float64)
        return X, np.clip(np.log((ps / pac).astype(f)), -Y_CLIP, Y_CLIP).astype(np.float64)
    X = np.stack([np.log(pr / pl), np.log(dr / dl), du / cl, du / cr, cl / (cl + cr),
                  cr / (cl + cr)], axis=1)
    return X, np.clip(np.log(ps / pac), -Y_CLIP, Y_CLIP)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[OFF_W1:OFF_B1].reshape(N_HIDDEN, N_IN), w[OFF_B1:OFF_W2],
            w[OFF_W2:OFF_B2].reshape(N_HIDDEN, N_HIDDEN), w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3],
            w[OFF_B3], w[OFF_MEAN:OFF_INV_STD], w[OFF_INV_STD:OFF_GATE])


def forward(w, X):
    """Network output y[n] and the activations needed by backward."""
    W1, b1, W2, b2, W3, b3, mean, inv_std = unpack(w)
    xn = (X - mean) * inv_std
    h1 = np.tanh(xn @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3, (xn, h1, h2)


def loss(w, X, Y):
    y, _ = forward(w, X)
    return float(np.mean((y - Y) ** 2))


def gradient(w, X, Y):
    """d loss / d (trained parameters), N_PARAMS values in block order."""
    W1, b1, W2, b2, W3, b3, mean, inv_std = unpack(w)
    y, (xn, h1, h2) = forward(w, X)
    g = 2.0 * (y - Y) / len(Y)                   # d loss / d y
    dW3 = h2.T @ g
    db3 = g.sum()
    d2 = (g[:, None] * W3[None, :]) * (1.0 - h2 ** 2)   # d loss / d a2
    dW2 = d2.T @ h1
    db2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * (1.0 - h1 ** 2)                   # d loss / d a1
    dW1 = d1.T @ xn
    db1 = d1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]])


def twoshock(kind, records, gamma, pmin):
    """p* and u* of every record's interface by the oracle's two-shock solver
    ("port": orc_twoshock, "ref": the reference's twoshock.F built by
    oracle/Makefile), called as the sweep driver calls it: slices of rows of
    ROW interfaces (twoshock.F:63-85 sizes its locals by
    MAX_ANY_SINGLE_DIRECTION, 1031), the last row padded."""
    ROW = 256
    r = np.asarray(records, dtype=np.float64)
    n = len(r)
    jdim = max(1, (n + ROW - 1) // ROW)
    pad = np.concatenate([r[:, :6], np.repeat(r[:1, :6], jdim * ROW - n, axis=0)])
    L = oracle.lib(kind)
    f = getattr(L, "orc_twoshock" if kind == "port" else "twoshock_")
    cols = [np.ascontiguousarray(pad[:, k]) for k in range(6)]
    pbar, ubar, gr = np.zeros(jdim * ROW), np.zeros(jdim * ROW), np.zeros(jdim * ROW)
    dp = C.POINTER(C.c_double)
    i64 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    f64 = lambda v: C.byref(C.c_double(v))  # noqa: E731
    f.restype = None
    f(*[c.ctypes.data_as(dp) for c in cols],
      i64(ROW), i64(jdim), i64(1), i64(ROW), i64(1), i64(jdim), f64(0.0), f64(gamma), f64(pmin),
      i64(0), pbar.ctypes.data_as(dp), ubar.ctypes.data_as(dp), i64(0), gr.ctypes.data_as(dp),
      i64(0), f64(0.0))
    return pbar[:n], ubar[:n]
