"""FP64 numpy oracle of the path-guiding MLP -- TEST INFRASTRUCTURE ONLY.

Same parameter layout, encoding and mixture parameterisation as the GPU
network (include/liar_amd.h, lr_guide_*).  Outputs 0..39 are the 8 vMF lobes,
40 and 41 the strategy-selection logits (light / guide / BSDF probabilities
of guided one-sample MIS), 42 the log of the reflected-radiance estimate
(radiance-driven Russian roulette), 43..47 padding.  With ``bf16=True`` the forward and
backward passes round to bfloat16 at the points where the GPU stores bf16
(weights, encoding, hidden activations, output and hidden gradients), so GPU
and oracle differ only by FP32-vs-FP64 accumulation.  The analytic gradient
is pinned by central finite differences in tests/test_guide_oracle.py.
"""
from __future__ import annotations

import numpy as np

W, HIDDEN, OUT, OUT_PAD, LOBES, FREQ, IN = 64, 5, 40, 48, 8, 4, 6
SEL, SEL_FLOOR = 2, 0.025
RAD, RAD_LOG_MAX, RAD_RATIO_MAX = 42, 20.0, 50.0
WT = HIDDEN * W * W + OUT_PAD * W
NPARAMS = WT + HIDDEN * W + OUT_PAD


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where((u & 0x7FFFFFFF) > 0x7F800000, 0x7FC00000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    Wt, b, o = [], [], 0
    for l in range(HIDDEN + 1):
        n = W if l < HIDDEN else OUT_PAD
        Wt.append(p[o:o + n * W].reshape(n, W))
        o += n * W
    for l in range(HIDDEN + 1):
        n = W if l < HIDDEN else OUT_PAD
        b.append(p[o:o + n])
        o += n
    return Wt, b


def kslot_of(f):
    """K slot (kb, j, e) of logical feature f of a layer's input (the GPU's
    MFMA operand layout): kb = f / 32, j = (f % 16) / 4, e = f % 4 (+4 past
    16 within the 32-block)."""
    kb, r = f >> 5, f & 31
    return kb, (r & 15) >> 2, (r & 3) + (4 if r >= 16 else 0)


def encode(inp):
    """64 features of the inputs' first 6 values, feature f at K slot (kb, j,
    e) = kslot_of(f), s = 8 kb + e: s < 12 sin (s even) / cos (s odd) of
    2^j pi v_{s/2}; s >= 12: j = 0 2 v_{s-12} - 1, j = 1 (2 v_4 - 1, 2 v_5 - 1,
    1, 0), j > 1 zero."""
    v = np.asarray(inp, np.float64)[:, :IN]
    n = v.shape[0]
    f = np.zeros((n, W))
    for c in range(W):
        kb, j, e = kslot_of(c)
        s = 8 * kb + e
        if s < 12:
            a = np.float32(np.pi) * (2 ** j) * v[:, s >> 1]
            f[:, c] = np.cos(a) if s & 1 else np.sin(a)
        elif j == 0:
            f[:, c] = 2 * v[:, s - 12] - 1
        elif j == 1:
            f[:, c] = 2 * v[:, s - 8] - 1 if s < 14 else (1.0 if s == 14 else 0.0)
    return f


def forward(p, inp, bf=False):
    Wt, b = unpack(p)
    if bf:
        Wt = [bf16(w) for w in Wt]
    a = encode(inp)
    if bf:
        a = bf16(a)
    acts = [a]
    for l in range(HIDDEN):
        z = acts[-1] @ Wt[l].T + b[l]
        a = np.maximum(z, 0)
        acts.append(bf16(a) if bf else a)
    out = acts[-1] @ Wt[HIDDEN].T + b[HIDDEN]
    return acts, out


KAPPA_LOG_MAX = 7.6


def kappa_of(s):
    """Lobe concentration 0.1 + e^min(s, 7.6)."""
    return 0.1 + np.exp(np.minimum(s, KAPPA_LOG_MAX))


def mixture(out):
    o = out[:, :OUT].reshape(-1, LOBES, 5)
    logit = o[..., 0]
    pi = np.exp(logit - logit.max(1, keepdims=True))
    pi /= pi.sum(1, keepdims=True)
    m = o[..., 1:4]
    mu = m / (np.linalg.norm(m, axis=2, keepdims=True) + 1e-12)
    kappa = kappa_of(o[..., 4])
    return pi, mu, kappa


def log_pdf(out, dirs):
    pi, mu, kappa = mixture(out)
    c = np.einsum("nkd,nd->nk", mu, dirs)
    lv = np.log(kappa) - np.log(2 * np.pi * (-np.expm1(-2 * kappa))) + kappa * (c - 1)
    lp = np.log(pi) + lv
    mx = lp.max(1, keepdims=True)
    return (mx + np.log(np.exp(lp - mx).sum(1, keepdims=True)))[:, 0]


def sigmoid(x):
    return 1 / (1 + np.exp(-x))


def selection(out):
    """(c_light, c_guide, c_bsdf) per row: floor + (1 - 3 floor) * (r_l,
    (1 - r_l) r_g, (1 - r_l)(1 - r_g)), r = sigmoid of outputs 40, 41."""
    rl, rg = sigmoid(out[:, OUT]), sigmoid(out[:, OUT + 1])
    ct = np.stack([rl, (1 - rl) * rg, (1 - rl) * (1 - rg)], 1)
    return SEL_FLOOR + (1 - 3 * SEL_FLOOR) * ct


def radiance(out):
    """Reflected-radiance estimate e^clamp(r, -20, 20) of output 42."""
    return np.exp(np.clip(out[:, RAD], -RAD_LOG_MAX, RAD_LOG_MAX))


def loss(p, inp, dirs, w, bf=False, pg=None, rad=0.0, rad_den=None):
    """sum_r w_r / sum(w) * (-log q(d_r) - [p_l + p_b > 0] log(c_l p_l + c_g
    p_g + c_b p_b)), with p_l, p_b = inp[:, 6], inp[:, 7], plus rad / n *
    sum_r (L_r - min(w_r, 50 D_r))^2 / D_r^2 with L the radiance estimate and
    D its value at the point of differentiation.  The guide pdf p_g in the
    selection term and D are constants of the optimisation (pass their values
    at that point as ``pg`` and ``rad_den``; default: the current ones)."""
    _, out = forward(p, inp, bf)
    wn = w / w.sum()
    lq = log_pdf(out, dirs)
    inp = np.asarray(inp, np.float64)
    pl, pb = inp[:, 6], inp[:, 7]
    on = (pl + pb) > 0
    c = selection(out)
    pgv = np.exp(lq) if pg is None else pg
    pmix = c[:, 0] * pl + c[:, 1] * pgv + c[:, 2] * pb
    sel = np.where(on, np.log(np.where(on, pmix, 1.0)), 0.0)
    total = -(wn * lq).sum() - (wn * sel).sum()
    if rad:
        L = radiance(out)
        D = L if rad_den is None else rad_den
        y = np.minimum(np.asarray(w, np.float64), RAD_RATIO_MAX * D)
        total += rad / len(L) * (((L - y) / D) ** 2).sum()
    return float(total)


def grad(p, inp, dirs, w, bf=False, rad=0.0):
    """Analytic gradient of loss() w.r.t. the flat parameter vector, except
    that the radiance term's gradient stops at output row 42 (the hidden
    layers get the gradient of loss(rad=0))."""
    Wt, b = unpack(p)
    Wtb = [bf16(x) for x in Wt] if bf else Wt
    acts, out = forward(p, inp, bf)
    n = out.shape[0]
    wn = w / w.sum()
    o = out[:, :OUT].reshape(n, LOBES, 5)
    pi, mu, kappa = mixture(out)
    c = np.einsum("nkd,nd->nk", mu, dirs)
    lv = np.log(kappa) - np.log(2 * np.pi * (-np.expm1(-2 * kappa))) + kappa * (c - 1)
    lp = np.log(pi) + lv
    logq = log_pdf(out, dirs)[:, None]
    resp = np.exp(lp - logq)
    g = np.zeros((n, OUT_PAD))
    gg = g[:, :OUT].reshape(n, LOBES, 5)
    gg[..., 0] = -wn[:, None] * (resp - pi)
    dlk = 1 / kappa - 2 / np.expm1(2 * kappa) + (c - 1)
    dk = np.where(o[..., 4] < KAPPA_LOG_MAX, np.exp(np.minimum(o[..., 4], KAPPA_LOG_MAX)), 0.0)
    gg[..., 4] = -wn[:, None] * resp * dlk * dk
    nm = np.linalg.norm(o[..., 1:4], axis=2) + 1e-12
    proj = dirs[:, None, :] - c[..., None] * mu
    gg[..., 1:4] = (-wn[:, None] * resp * kappa / nm)[..., None] * proj
    g[:, :OUT] = gg.reshape(n, OUT)
    inp = np.asarray(inp, np.float64)
    pl, pb = inp[:, 6], inp[:, 7]
    on = (pl + pb) > 0
    pg = np.exp(logq[:, 0])
    c = selection(out)
    pmix = np.where(on, c[:, 0] * pl + c[:, 1] * pg + c[:, 2] * pb, 1.0)
    rl, rg = sigmoid(out[:, OUT]), sigmoid(out[:, OUT + 1])
    f = np.where(on, -wn * (1 - 3 * SEL_FLOOR) / pmix, 0.0)
    g[:, OUT] = f * rl * (1 - rl) * (pl - rg * pg - (1 - rg) * pb)
    g[:, OUT + 1] = f * (1 - rl) * rg * (1 - rg) * (pg - pb)
    if rad:
        r = out[:, RAD]
        ratio = np.minimum(np.asarray(w, np.float64) / radiance(out), RAD_RATIO_MAX)
        inside = (r > -RAD_LOG_MAX) & (r < RAD_LOG_MAX)
        g[:, RAD] = np.where(inside, rad / n * 2 * (1 - ratio), 0.0)
    if bf:
        g = bf16(g)
    dW = [None] * (HIDDEN + 1)
    db = [None] * (HIDDEN + 1)
    dz = g
    for l in range(HIDDEN, -1, -1):
        dW[l] = dz.T @ acts[l]
        db[l] = dz.sum(0)
        if l == 0:
            break
        if l == HIDDEN:
            # output 42 is a readout of the last hidden layer: its loss does
            # not reach the layers the mixture and the selection share
            dz = dz.copy()
            dz[:, RAD] = 0.0
        da = dz @ Wtb[l]
        dz = np.where(acts[l] > 0, da, 0.0)
        if bf:
            dz = bf16(dz)
    return np.concatenate([x.ravel() for x in dW] + [x.ravel() for x in db])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    Wt, _ = unpack(p)
    o = 0
    for l in range(HIDDEN + 1):
        nrow = W if l < HIDDEN else OUT_PAD
        lim = np.sqrt(6 / W) * (1 if l < HIDDEN else 0.1)
        blk = rng.uniform(-lim, lim, (nrow, W))
        if l == HIDDEN:
            blk[OUT + SEL:] = 0
        p[o:o + nrow * W] = blk.ravel()
        o += nrow * W
    return p
