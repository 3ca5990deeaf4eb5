"""Pins the FP64 guiding-network oracle (oracle/guide.py): analytic gradient
vs central finite differences (lobes, selection and the reflected-radiance
term, whose gradient stops at its readout row), the radiance term's
stationary point at the mean target, the encoding's feature set and slot
order, the selection optimum against a brute-force grid, mixture pdf
normalised on the sphere, bf16 rounding matches the IEEE
round-to-nearest-even definition."""
import numpy as np

from oracle import guide as G


def _case(seed=0, n=6):
    rng = np.random.default_rng(seed)
    p = G.init_params(seed + 1)
    p[G.WT:] = rng.normal(0, 0.1, p.size - G.WT)
    inp = rng.random((n, 8))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.random(n) + 0.5
    return p, inp, d, w


def test_gradient_matches_finite_differences():
    p, inp, d, w = _case()
    g = G.grad(p, inp, d, w)
    rng = np.random.default_rng(3)
    # weights of every layer, biases, including the output layer
    idx = np.concatenate([rng.choice(G.WT, 48, replace=False), G.WT + rng.choice(p.size - G.WT, 16, replace=False)])
    # the selection rows (outputs 40, 41) and their biases, always
    out0 = G.HIDDEN * G.W * G.W
    idx = np.concatenate([idx, out0 + G.OUT * G.W + rng.choice(2 * G.W, 8, replace=False),
                          [G.WT + G.HIDDEN * G.W + G.OUT, G.WT + G.HIDDEN * G.W + G.OUT + 1]])
    # the guide pdf inside the selection term is held at its value here
    _, out = G.forward(p, inp)
    pg = np.exp(G.log_pdf(out, d))
    fd = []
    for k in idx:
        e = np.zeros(p.size)
        e[k] = 1e-6
        fd.append((G.loss(p + e, inp, d, w, pg=pg) - G.loss(p - e, inp, d, w, pg=pg)) / 2e-6)
    fd = np.array(fd)
    assert np.max(np.abs(fd - g[idx])) < 1e-6 * max(1.0, np.abs(g[idx]).max())
    sel_rows = g[out0 + G.OUT * G.W:out0 + (G.OUT + 2) * G.W]
    assert np.abs(sel_rows).max() > 1e-4  # the selection term is exercised
    # padded output rows 42..47 get no gradient
    out_rows = g[out0:G.WT].reshape(G.OUT_PAD, G.W)
    assert (out_rows[G.OUT + G.SEL:] == 0).all()
    # rows without light / BSDF pdfs (inputs 6, 7 zero) train no selection
    inp0 = inp.copy()
    inp0[:, 6:8] = 0
    g0 = G.grad(p, inp0, d, w)
    assert (g0[out0 + G.OUT * G.W:out0 + (G.OUT + 2) * G.W] == 0).all()


def test_radiance_gradient_matches_finite_differences():
    """The reflected-radiance term (output 42, relative loss with its
    denominator held fixed, targets clipped at 50x the estimate) against
    central differences; one row's target is past the clip.  The term trains
    the readout row only."""
    p, inp, d, w = _case(seed=4, n=7)
    w[2] = 400.0  # w / L > 50: the clipped branch
    rad = 0.7
    g = G.grad(p, inp, d, w, rad=rad)
    out0 = G.HIDDEN * G.W * G.W
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(out0, 24, replace=False),
                          out0 + G.RAD * G.W + rng.choice(G.W, 12, replace=False),
                          [G.WT + G.HIDDEN * G.W + G.RAD], G.WT + rng.choice(G.HIDDEN * G.W, 8, replace=False)])
    _, out = G.forward(p, inp)
    pg = np.exp(G.log_pdf(out, d))
    D = G.radiance(out)
    assert (np.minimum(w, 50 * D) < w).any()
    # output row 42 and its bias: the whole loss; the hidden layers: the loss
    # without the radiance term (its gradient stops at the readout row)
    row42 = set(range(out0 + G.RAD * G.W, out0 + (G.RAD + 1) * G.W)) | {G.WT + G.HIDDEN * G.W + G.RAD}
    fd = []
    for k in idx:
        e = np.zeros(p.size)
        e[k] = 1e-6
        r = rad if k in row42 else 0.0
        fd.append((G.loss(p + e, inp, d, w, pg=pg, rad=r, rad_den=D)
                   - G.loss(p - e, inp, d, w, pg=pg, rad=r, rad_den=D)) / 2e-6)
    fd = np.array(fd)
    assert np.max(np.abs(fd - g[idx])) < 1e-6 * max(1.0, np.abs(g[idx]).max())
    assert np.abs(g[out0 + G.RAD * G.W:out0 + (G.RAD + 1) * G.W]).max() > 1e-3
    # without the term output 42 gets no gradient, and every other parameter
    # the same gradient as with it
    g0 = G.grad(p, inp, d, w)
    assert (g0[out0 + G.RAD * G.W:out0 + (G.RAD + 1) * G.W] == 0).all()
    other = np.ones(p.size, bool)
    other[sorted(row42)] = False
    assert np.array_equal(g0[other], g[other])


def test_radiance_stationary_point_is_the_mean():
    """With output 42 a bias alone, the radiance gradient vanishes exactly
    where e^b is the mean target and pushes b towards it from either side
    (a log-space or unweighted-ratio loss would settle elsewhere)."""
    p, inp, d, _ = _case(seed=6, n=50)
    rng = np.random.default_rng(7)
    w = rng.exponential(2.5, 50) * (rng.random(50) < 0.6)  # many zero targets
    out0 = G.HIDDEN * G.W * G.W
    p[out0 + G.RAD * G.W:out0 + (G.RAD + 1) * G.W] = 0.0
    bi = G.WT + G.HIDDEN * G.W + G.RAD
    grads = []
    for off in (0.0, 0.2, -0.2):
        p[bi] = np.log(w.mean()) + off
        grads.append(G.grad(p, inp, d, w, rad=1.0)[bi])
    assert abs(grads[0]) < 1e-12 and grads[1] > 1e-3 and grads[2] < -1e-3


def test_encoding_features():
    """Every row's 64 features are, in some order, the 6 inputs as 2v - 1,
    sin and cos of 2^k pi v_d for k < 4 and every input d, one constant 1
    and 9 zeros; feature f's position is the GPU's K slot of f (the order the
    kernels compute them in), checked on three hand-placed slots."""
    rng = np.random.default_rng(11)
    v = rng.random((5, 8))
    f = G.encode(v)
    for r in range(5):
        want = [2 * v[r, d] - 1 for d in range(6)] + [1.0] + [0.0] * 9
        for d in range(6):
            for k in range(4):
                a = np.float32(np.pi) * 2 ** k * v[r, d]
                want += [np.sin(a), np.cos(a)]
        assert np.allclose(np.sort(f[r]), np.sort(want), atol=1e-15)
    # f = 11 is slot (kb 0, j 2, e 3): s = 3, cos(4 pi v_1); f = 54 is (1, 1, 6):
    # s = 14 with j = 1, the constant; f = 1 is (0, 0, 1): s = 1, cos(pi v_0)
    assert np.allclose(f[:, 4 * 2 + 3], np.cos(np.float32(np.pi) * 4 * v[:, 1]))
    assert (f[:, 32 + 16 + 4 * 1 + 2] == 1.0).all()
    assert np.allclose(f[:, 1], np.cos(np.float32(np.pi) * v[:, 0]))


def test_selection_optimum_matches_closed_form():
    # With the target concentrated where only one strategy has density, the
    # selection loss drives that strategy's probability to its maximum
    # 1 - 2 floor.  Brute force over the two logits on a grid.
    rng = np.random.default_rng(9)
    n = 400
    pl = np.where(rng.random(n) < 0.8, 5.0, 0.0)   # light covers 80 % of the records
    pb = np.full(n, 0.3)                           # BSDF covers all, thinly
    pg = np.full(n, 0.01)
    best = None
    for a0 in np.linspace(-6, 6, 61):
        for a1 in np.linspace(-6, 6, 61):
            out = np.zeros((1, G.OUT_PAD))
            out[0, G.OUT], out[0, G.OUT + 1] = a0, a1
            c = G.selection(out)[0]
            val = -np.log(c[0] * pl + c[1] * pg + c[2] * pb).mean()
            if best is None or val < best[0]:
                best = (val, c)
    c = best[1]
    assert abs(c.sum() - 1) < 1e-12 and (c >= G.SEL_FLOOR - 1e-12).all()
    # light share close to the 0.8 of records only it explains well; BSDF takes the rest
    assert 0.7 < c[0] < 0.9 and c[2] > c[1]


def test_mixture_pdf_integrates_to_one():
    p, inp, _, _ = _case(seed=4, n=3)
    p[G.HIDDEN * G.W * G.W:G.WT] *= 20  # sharper lobes
    _, out = G.forward(p, inp)
    # quadrature on a fine (z, phi) grid: dA = dz dphi
    nz, nphi = 800, 400
    z = -1 + (np.arange(nz) + 0.5) * 2 / nz
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    Z, P = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - Z ** 2)
    dirs = np.stack([r * np.cos(P), r * np.sin(P), Z], -1).reshape(-1, 3)
    for row in range(3):
        lq = G.log_pdf(np.repeat(out[row:row + 1], len(dirs), 0), dirs)
        total = np.exp(lq).sum() * (2 / nz) * (2 * np.pi / nphi)
        assert abs(total - 1) < 2e-3, total


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-39, 65504.0, np.inf], np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even (1 + 2^-8 -> 1)
    assert r[2] == 1.0078125            # above the tie rounds up
    assert r[4] == 65536.0 and np.isinf(r[5])
