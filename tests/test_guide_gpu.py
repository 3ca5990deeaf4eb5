"""GPU guiding network (BF16 MFMA) against the FP64 oracle of oracle/guide.py.

- forward: GPU mixture parameters equal the oracle's with bf16 rounding at the
  same points (differences: FP32 vs FP64 accumulation only); a wrong MFMA
  lane layout or layer wiring fails this;
- gradient: lr_guide_train_step's gradient (learning rate 0) equals the
  oracle's analytic gradient, itself pinned by finite differences;
- training: Adam on samples of a known position-dependent vMF target lowers
  the loss from ~log(4 pi) to near the target's entropy; the reflected-
  radiance output learns the conditional mean of noisy targets.
"""
import numpy as np
import pytest

from liar_amd.guide import Guide
from oracle import guide as G

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 8), np.float32)
    x[:, :6] = rng.random((n, 6))
    return x


def test_forward_matches_oracle(gpu_ctx):
    import torch

    g = Guide(gpu_ctx, seed=3, learning_rate=0.0)
    p = g.params().astype(np.float64)
    rng = np.random.default_rng(1)
    p[G.WT:] = rng.normal(0, 0.05, p.size - G.WT)  # nonzero biases
    p[G.HIDDEN * G.W * G.W:G.WT] *= 8  # informative outputs
    g.set_params(p.astype(np.float32))
    p = g.params().astype(np.float64)
    for n in (1, 17, 1000):
        x = _inputs(n, n)
        mix = g.infer(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
        _, out = G.forward(p, x.astype(np.float64), bf=True)
        pi, mu, kappa = G.mixture(out)
        got = mix[:, :40].reshape(n, 8, 5)
        e_c = np.abs(mix[:, 40:43] - G.selection(out)).max()
        e_r = np.abs(mix[:, 43] / G.radiance(out) - 1).max()
        assert e_c < 5e-3 and e_r < 1e-2 and (mix[:, 44:] == 0).all()
        # single-ulp bf16 rounding flips (FP32 vs FP64 accumulation) move
        # values by ~1e-3; a wrong lane layout or wiring moves them by O(0.1)
        e_pi, e_mu = np.abs(got[..., 0] - pi).max(), np.abs(got[..., 1:4] - mu).max()
        e_k = (np.abs(got[..., 4] - kappa) / kappa).max()
        print(f"[gpu] guide forward n={n}: max |d pi| {e_pi:.1e}, |d mu| {e_mu:.1e}, rel d kappa {e_k:.1e}")
        assert e_pi < 5e-3 and e_mu < 1e-2 and e_k < 1e-2
        assert np.abs(got[..., 0] - pi).mean() < 2e-4
    g.close()


@pytest.mark.parametrize("n,rad", [(3000, 0.0), (40000, 0.0), (3000, 0.7), (40000, 0.7)])
def test_gradient_matches_oracle(gpu_ctx, n, rad):
    """n = 40000 spans several row groups per workgroup (the weight gradients
    accumulate across them) and ends in a partial group; rad > 0 adds the
    reflected-radiance loss."""
    import torch

    g = Guide(gpu_ctx, seed=5, learning_rate=0.0, radiance_weight=rad)
    p = g.params().astype(np.float64)
    rng = np.random.default_rng(2)
    p[G.WT:] = rng.normal(0, 0.05, p.size - G.WT)
    p[G.HIDDEN * G.W * G.W:G.WT] *= 6
    g.set_params(p.astype(np.float32))
    p = g.params().astype(np.float64)
    x = _inputs(n, 7)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = (rng.random(n) + 0.2).astype(np.float32)
    # light / BSDF pdfs of the records (zero for a third: no selection term there)
    x[:, 6] = np.where(rng.random(n) < 0.5, rng.random(n) * 4, 0.0)
    x[:, 7] = np.where(np.arange(n) % 3 == 0, 0.0, rng.random(n) * 0.6)
    x[np.arange(n) % 3 == 0, 6] = 0.0
    grad = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    g.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(d.astype(np.float32)).cuda(),
                 torch.from_numpy(w).cuda(), loss=loss, grad=grad)
    torch.cuda.synchronize()
    gg = grad.cpu().numpy().astype(np.float64)
    go = G.grad(p, x.astype(np.float64), d, w.astype(np.float64), bf=True, rad=rad)
    rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
    lo = G.loss(p, x.astype(np.float64), d, w.astype(np.float64), bf=True, rad=rad)
    print(f"[gpu] guide gradient n={n} rad={rad} rel err {rel:.2e}, loss gpu {loss.item():.6f} oracle {lo:.6f}")
    assert rel < 1e-2
    # every layer's weight and bias block on its own (a wrong tile or row
    # mapping in one small block would hide in the global norm)
    for layer in range(G.HIDDEN + 1):
        nout = G.W if layer < G.HIDDEN else 48
        blocks = [slice(layer * G.W * G.W, layer * G.W * G.W + nout * G.W),
                  slice(G.WT + layer * G.W, G.WT + layer * G.W + nout)]
        for sl in blocks:
            e = np.linalg.norm(gg[sl] - go[sl]) / max(np.linalg.norm(go[sl]), 1e-30)
            assert e < 3e-2, (layer, sl, e)
    out0 = G.HIDDEN * G.W * G.W
    row42 = slice(out0 + G.RAD * G.W, out0 + (G.RAD + 1) * G.W)
    if rad:  # the radiance row on its own
        assert np.linalg.norm(go[row42]) > 0
        assert np.linalg.norm(gg[row42] - go[row42]) / np.linalg.norm(go[row42]) < 3e-2
    else:
        assert (gg[row42] == 0).all()
    assert abs(loss.item() - lo) < 1e-3 * max(1, abs(lo))
    assert np.array_equal(g.params(), p.astype(np.float32))  # learning rate 0: no update
    g.close()


def _sample_vmf(mu, kappa, rng):
    n = len(mu)
    xi = rng.random(n)
    wv = 1 + np.log(xi + (1 - xi) * np.exp(-2 * kappa)) / kappa
    phi = 2 * np.pi * rng.random(n)
    t = np.cross(mu, np.where(np.abs(mu[:, :1]) < 0.9, [[1, 0, 0]], [[0, 1, 0]]))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(mu, t)
    s = np.sqrt(np.maximum(0, 1 - wv ** 2))
    return wv[:, None] * mu + s[:, None] * (np.cos(phi)[:, None] * t + np.sin(phi)[:, None] * b)


def test_training_learns_radiance_mean(gpu_ctx):
    """The radiance output is a linear readout of the last hidden layer (its
    loss does not train the shared layers), so it learns what the features the
    lobes need can express: with the vMF target of the test below (its mean
    direction a function of the position) and row weights y = m(x) * E, E ~
    Exp(1) on 70% of the rows and 0 on the rest, m(x) = 0.2 + 3 x0^2, it
    converges to the conditional mean 0.7 m(x), not to a log-space or median
    fit (those sit at 0.48 m(x) and lower)."""
    import torch

    g = Guide(gpu_ctx, seed=12, learning_rate=3e-3, radiance_weight=1.0)
    rng = np.random.default_rng(8)
    for step in range(600):
        n = 8192
        x = _inputs(n, 5000 + step)
        mu = x[:, :3] - 0.5 + np.array([0, 0, 0.3])
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        d = _sample_vmf(mu.astype(np.float64), 20.0, rng).astype(np.float32)
        m = 0.2 + 3.0 * x[:, 0] ** 2
        y = (m * rng.exponential(1.0, n) * (rng.random(n) < 0.7)).astype(np.float32)
        g.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(d).cuda(), torch.from_numpy(y).cuda())
    x = _inputs(4096, 99)
    mix = g.infer(torch.from_numpy(x).cuda()).cpu().numpy()
    want = 0.7 * (0.2 + 3.0 * x[:, 0] ** 2)
    rel = np.abs(mix[:, 43] / want - 1)
    print(f"[gpu] radiance readout: mean rel err {rel.mean():.3f}, 95th pct {np.quantile(rel, 0.95):.3f}, "
          f"mean ratio {np.mean(mix[:, 43] / want):.3f}")
    assert rel.mean() < 0.15 and abs(np.mean(mix[:, 43] / want) - 1) < 0.1
    g.close()


def test_training_learns_a_vmf_target(gpu_ctx):
    import torch

    g = Guide(gpu_ctx, seed=11, learning_rate=3e-3)
    rng = np.random.default_rng(5)
    kappa = 20.0
    losses = []
    for step in range(600):
        n = 8192
        x = _inputs(n, 1000 + step)
        mu = x[:, :3] - 0.5 + np.array([0, 0, 0.3])
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        d = _sample_vmf(mu.astype(np.float64), kappa, rng).astype(np.float32)
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        g.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(d).cuda(),
                     torch.ones(n, dtype=torch.float32, device="cuda"), loss=loss)
        losses.append(loss.item())
    entropy = -np.log(kappa / (2 * np.pi * (1 - np.exp(-2 * kappa)))) + 1 - kappa * (
        1 / np.tanh(kappa) - 1 / kappa - (1 - 1 / kappa))
    print(f"[gpu] guide training: loss {losses[0]:.3f} -> {np.mean(losses[-20:]):.3f} "
          f"(uniform {np.log(4 * np.pi):.3f}, target entropy {entropy:.3f})")
    assert losses[0] > 2.0
    assert np.mean(losses[-20:]) < 0.6
    g.close()
