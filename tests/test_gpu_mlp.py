"""The warm-start network's matrix-core kernels (BF16 operands, FP32 accumulation) against the
CPU FP64 reference (tests/mlp_ref.py) that rounds to bfloat16 at the same points: forward
outputs, the gradient of the mean squared error, and training that lowers the loss."""
import sys

import numpy as np
import pytest

import mlp_ref as R

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.__stderr__, flush=True)


def _setup(seed=0, n=1000):
    import torch

    from wflowhip import Mlp

    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.25, R.NPARAM).astype(np.float32)
    x = rng.normal(2.0, 3.0, (n, R.NIN)).astype(np.float32)
    xm = x.mean(0).astype(np.float32)
    xs = x.std(0).astype(np.float32)
    t = rng.normal(0.0, 1.0, n).astype(np.float32)
    net = Mlp(device=0, seed=seed)
    net.set_params(p)
    net.set_normalization(xm, xs, 0.0, 1.0)
    xd = torch.as_tensor(x, device="cuda").contiguous()
    td = torch.as_tensor(t, device="cuda").contiguous()
    return net, p, x, xm, xs, t, xd, td


def test_forward_matches_fp64_reference():
    import torch

    net, p, x, xm, xs, t, xd, td = _setup(1, 1000)  # 1000: a partial last wave
    y = net.predict(xd, torch.empty(len(t), dtype=torch.float32, device="cuda")).cpu().numpy()
    xn = R.features_in(x, xm, xs)
    yr, _ = R.forward(p.astype(np.float64), xn)
    err = np.abs(y - yr) / (np.abs(yr) + 1e-2 * np.abs(yr).mean())
    _say(f"forward: max rel err {err.max():.2e} over {len(t)} samples")
    assert np.median(err) < 1e-5 and np.quantile(err, 0.99) < 1e-3, (np.median(err), err.max())
    net.close()


def test_gradient_matches_fp64_reference():
    net, p, x, xm, xs, t, xd, td = _setup(2, 4096)
    g, loss = net.gradient(xd, td)
    gr, lr_ = R.gradient(p.astype(np.float64), R.features_in(x, xm, xs), t.astype(np.float64))
    assert abs(loss - lr_) <= 1e-5 * lr_
    blocks = {"W1": (R.OW1, R.OB1), "b1": (R.OB1, R.OW2), "W2": (R.OW2, R.OB2), "b2": (R.OB2, R.OW3),
              "w3": (R.OW3, R.OB3), "b3": (R.OB3, R.NPARAM)}
    for name, (a, b) in blocks.items():
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        _say(f"gradient {name}: relative L2 error {rel:.2e}")
        assert rel < 2e-3, (name, rel)
    net.close()


def test_training_lowers_the_loss():
    import torch

    from wflowhip import Mlp

    rng = np.random.default_rng(4)
    n = 65536
    x = rng.normal(0.0, 1.0, (n, R.NIN)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2).astype(np.float32)
    net = Mlp(device=0, seed=7)
    net.set_normalization(np.zeros(8, np.float32), np.ones(8, np.float32), 0.0, 1.0)
    xd = torch.as_tensor(x, device="cuda").contiguous()
    td = torch.as_tensor(t, device="cuda").contiguous()
    losses = [net.train_epoch(xd, td, batch=4096, lr=3e-3) for _ in range(8)]
    _say(f"training losses {['%.4f' % v for v in losses]}")
    assert losses[-1] < 0.25 * losses[0], losses
    net.close()
