"""numpy (float64) restatement of the learned per-sector preconditioner's
training pass (test helper): features of each sector block, the MLP
8 -> 256 (tanh) -> 256 (tanh) -> 2, the head sigma = SIGMA0 scale
exp(BOUND tanh(o0)), alpha = exp(BOUND tanh(o1)) (a bounded correction of the
fixed diagonal form sigma = SIGMA0 scale, alpha = 1), the preconditioner
t = alpha r / (max(d - theta, 0) + sigma), the loss
1 - Re<t, delta>^2 / (|t|^2 |delta|^2) and its gradient with respect to every network parameter.  The GPU kernels
(hh_precond.hip) are checked against it; it is checked against central
finite differences (tests/test_precond_ref.py)."""
import numpy as np

NF, H, NO = 8, 256, 2
SIGMA0 = 0.1
BOUND = np.log(4.0)
W1 = 0
B1 = W1 + H * NF
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + NO * H
SIZE = B3 + NO


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[W1:B1].reshape(H, NF), w[B1:W2], w[W2:B2].reshape(H, H), w[B2:W3], w[W3:B3].reshape(NO, H),
            w[B3:SIZE])


def blocks(rows, cols, hs, he):
    """Per sector: element offset, and d in column-major order (a fast)."""
    d, off, so, eo, o = [], [], 0, 0, 0
    for R, Cc in zip(rows, cols):
        dd = hs[so:so + R][:, None] + he[eo:eo + Cc][None, :]
        d.append(dd.reshape(-1, order="F"))
        off.append(o)
        so += R
        eo += Cc
        o += R * Cc
    return off, np.concatenate(d)


def features(rows, cols, s2, nq, n_mid, hs, he, r, theta):
    off, d = blocks(rows, cols, hs, he)
    cnt = np.array([R * Cc for R, Cc in zip(rows, cols)], dtype=np.float64)
    mean_g = d.mean()
    sd = np.sqrt(max((d * d).mean() - mean_g * mean_g, 0.0))
    sd = max(sd, max(1e-3 * abs(mean_g), 1e-12))
    rr_tot = float(np.sum(np.abs(r) ** 2))
    lr = 0.5 * np.log10(max(rr_tot, 1e-300))
    F = np.zeros((len(rows), NF))
    for p, (o, c) in enumerate(zip(off, cnt)):
        dp = d[o:o + int(c)]
        rp = r[o:o + int(c)]
        m = dp.mean()
        v = max((dp * dp).mean() - m * m, 0.0)
        F[p] = [(m - theta) / sd, (dp.min() - theta) / sd, np.sqrt(v) / sd, np.log(c) / 10.0,
                np.log(np.sum(np.abs(rp) ** 2) / max(rr_tot, 1e-300) + 1e-30) / 10.0, lr / 4.0, s2[p] / 8.0,
                (nq[p] - n_mid) / 8.0]
    return off, d, sd, F


def forward(F, w):
    w1, b1, w2, b2, w3, b3 = unpack(w)
    h1 = np.tanh(F @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    out = h2 @ w3.T + b3
    return h1, h2, out


def head(out, sd):
    return SIGMA0 * sd * np.exp(BOUND * np.tanh(out[:, 0])), np.exp(BOUND * np.tanh(out[:, 1]))


def loss_and_grad(rows, cols, s2, nq, n_mid, hs, he, r, delta, theta, w, bug=None):
    """(loss, dL/dw, features, outputs, sigma, alpha).  bug plants a mistake
    in the backward pass ("no_tanh_grad": the second layer's tanh derivative
    left out; "w2_transposed": the second layer's weight transposed) so that
    the finite-difference check can be shown to catch it."""
    off, d, sd, F = features(rows, cols, s2, nq, n_mid, hs, he, r, theta)
    w1, b1, w2, b2, w3, b3 = unpack(w)
    h1, h2, out = forward(F, w)
    sig, al = head(out, sd)
    dn = float(np.sum(np.abs(delta) ** 2))
    t = np.zeros_like(r)
    den = np.zeros(len(r))
    for p, o in enumerate(off):
        c = rows[p] * cols[p]
        den[o:o + c] = np.maximum(d[o:o + c] - theta, 0.0) + sig[p]
        t[o:o + c] = al[p] * r[o:o + c] / den[o:o + c]
    A = float(np.sum(np.real(np.conj(delta) * t)))
    B = float(np.sum(np.abs(t) ** 2))
    L = 1.0 - A * A / (B * dn)
    dout = np.zeros_like(out)
    for p, o in enumerate(off):
        c = rows[p] * cols[p]
        tp, dp, nn = t[o:o + c], delta[o:o + c], den[o:o + c]
        dA_ds = -np.sum(np.real(np.conj(dp) * tp) / nn)
        dB_ds = -2.0 * np.sum(np.abs(tp) ** 2 / nn)
        dA_da = np.sum(np.real(np.conj(dp) * tp)) / al[p]
        dB_da = 2.0 * np.sum(np.abs(tp) ** 2) / al[p]
        gs = -2.0 * A * dA_ds / (B * dn) + A * A * dB_ds / (B * B * dn)
        ga = -2.0 * A * dA_da / (B * dn) + A * A * dB_da / (B * B * dn)
        dout[p, 0] = gs * sig[p] * BOUND * (1.0 - np.tanh(out[p, 0]) ** 2)
        dout[p, 1] = ga * al[p] * BOUND * (1.0 - np.tanh(out[p, 1]) ** 2)
    g = np.zeros(SIZE)
    gw3 = dout.T @ h2
    gb3 = dout.sum(0)
    dh2 = (dout @ w3) * (1.0 if bug == "no_tanh_grad" else (1.0 - h2 ** 2))
    gw2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ (w2.T if bug == "w2_transposed" else w2)) * (1.0 - h1 ** 2)
    gw1 = dh1.T @ F
    gb1 = dh1.sum(0)
    g[W1:B1] = gw1.reshape(-1)
    g[B1:W2] = gb1
    g[W2:B2] = gw2.reshape(-1)
    g[B2:W3] = gb2
    g[W3:B3] = gw3.reshape(-1)
    g[B3:SIZE] = gb3
    return L, g, F, out, sig, al


def loss_only(rows, cols, s2, nq, n_mid, hs, he, r, delta, theta, w):
    return loss_and_grad(rows, cols, s2, nq, n_mid, hs, he, r, delta, theta, w)[0]


def random_problem(seed, np_=5):
    """A small random sector table, diagonal, residual, target and network."""
    rng = np.random.default_rng(seed)
    rows = rng.integers(2, 9, np_).astype(np.int32)
    cols = rng.integers(2, 9, np_).astype(np.int32)
    s2 = rng.integers(0, 5, np_).astype(np.int32)
    nq = rng.integers(10, 20, np_).astype(np.int32)
    hs = rng.normal(2.0, 1.0, int(rows.sum()))
    he = rng.normal(1.0, 1.0, int(cols.sum()))
    n = int(np.sum(rows * cols))
    r = rng.normal(size=n) + 1j * rng.normal(size=n)
    delta = 0.3 * r / (np.abs(rng.normal(2.0, 0.5, n)) + 0.5) + 0.05 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    w = np.zeros(SIZE)
    w[W1:B1] = rng.uniform(-1, 1, H * NF) * np.sqrt(6.0 / (NF + H))
    w[B1:W2] = rng.normal(0, 0.1, H)
    w[W2:B2] = rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / (2 * H))
    w[B2:W3] = rng.normal(0, 0.1, H)
    w[W3:B3] = rng.uniform(-1, 1, NO * H) * 0.1
    w[B3:SIZE] = [0.3, 0.1]
    return rows, cols, s2, nq, 12, hs, he, r, delta, 0.5, w.astype(np.float32).astype(np.float64)
