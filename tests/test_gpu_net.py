"""The learned preconditioner's network on the GPU (BF16 MFMA forward,
data-gradient and weight-gradient kernels, FP32 accumulation) against the
CPU network oracle with the same BF16 operand rounding, and training on
the GPU.  Gates: forward relative L2 5e-3, loss 1e-3, gradient relative L2
1e-2 (the oracle accumulates in FP64, the MFMA in FP32; a BF16 rounding of
an activation may land on the other side of a tie)."""
import numpy as np
import pytest

import test_oracle_net as tn

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("variant", [1, 0])
@pytest.mark.parametrize("H,W,B", [(16, 8, 2), (24, 72, 3)])
def test_forward_and_gradients_match_oracle(H, W, B, variant):
    import snac_amd
    print(f"net check H={H} W={W} B={B} conv variant {variant}", flush=True)
    snac_amd.Net.conv_variant(variant)
    p, x, t = tn.random_problem(B=B, H=H, W=W, seed=H + W)
    net = snac_amd.Net(H, W, maxbatch=4)
    net.set_params(p)
    y = net.forward(x)
    yo = np.zeros_like(y)
    assert tn.lib().on_forward(B, H, W, p, x, yo, 1) == 0
    assert rel(y, yo) <= 5e-3, rel(y, yo)
    loss, g = net.loss_grad(x, t)
    lo, go = tn.loss_grad(p, x, t, round_=1)
    assert abs(loss - lo) <= 1e-3 * lo
    assert rel(g, go) <= 1e-2, rel(g, go)
    net.close()
    snac_amd.Net.conv_variant(1)


def test_training_lowers_the_loss():
    """Learn the solution operator of a periodic 2D Helmholtz problem
    (-Lap + lambda) u = f on 32 x 32 planes, lambda given as the second
    input channel; 20 epochs of Adam over 64 samples halve the loss."""
    import snac_amd
    H = W = 32
    n = 64
    rng = np.random.default_rng(7)
    kx = 2 * np.pi * np.fft.fftfreq(W)
    ky = 2 * np.pi * np.fft.fftfreq(H)
    sym = 4 * np.sin(ky[:, None] / 2) ** 2 + 4 * np.sin(kx[None, :] / 2) ** 2
    f = rng.standard_normal((n, H, W))
    lam = rng.uniform(0.05, 2.0, n)
    u = np.real(np.fft.ifft2(np.fft.fft2(f) / (sym[None] + lam[:, None, None])))
    x = np.stack([f / f.std(axis=(1, 2), keepdims=True), np.broadcast_to(lam[:, None, None], f.shape)], axis=-1)
    t = u / f.std(axis=(1, 2), keepdims=True)
    net = snac_amd.Net(H, W, maxbatch=8, seed=3)
    hist = net.train(x.astype(np.float32), t.astype(np.float32), batch=8, epochs=20, lr=2e-3)
    print("loss per epoch", hist, flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.5 * hist[0]
    net.close()
