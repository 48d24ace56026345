"""CPU FP64 reference of the warm-start network (8-64-64-1) and its mean-squared-error gradient,
for checking the library's matrix-core forward and backward kernels (test infrastructure).

`forward` / `gradient` with bf16=True round to bfloat16 at the points where the kernels hand a
value to a matrix-core product (normalised features, W1, W2, h1, and the output gradients of
layers 2 and 1 in the weight-gradient products); every sum is FP64. With bf16=False the network
is plain FP64, which `tests/test_mlp_cpu.py` pins by central finite differences.
"""
import numpy as np

NIN, NH = 8, 64
OW1, OB1 = 0, NH * NIN
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NPARAM = OB3 + 1


def bf16(x):
    """float32 -> bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32)
    r = (a + (((a >> 16) & 1) + 0x7FFF)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(NH, NIN), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3], p[OW3:OB3], p[OB3])


def features_in(x, x_mean, x_std, bf=True):
    """Normalised features as the kernels form them: (x - mean) * (1 / std) in float32."""
    x = np.asarray(x, np.float32)
    inv = (np.float32(1.0) / np.asarray(x_std, np.float32)).astype(np.float32)
    xn = ((x - np.asarray(x_mean, np.float32)) * inv).astype(np.float32)
    return bf16(xn) if bf else xn.astype(np.float64)


def forward(p, xn, bf=True):
    """xn: [n, 8] normalised features -> y [n] and the tape."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    z1 = xn @ r(W1).T + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ r(W2).T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return y, (xn, z1, h1, z2, h2)


def gradient(p, xn, t, bf=True):
    """Gradient of mean((y - t)^2) over the n samples, in the flat parameter order."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    y, (xn, z1, h1, z2, h2) = forward(p, xn, bf)
    n = len(t)
    dy = 2.0 * (y - np.asarray(t, np.float64)) / n
    g = np.zeros(NPARAM)
    g[OW3:OB3] = h2.T @ dy
    g[OB3] = dy.sum()
    g2 = (z2 > 0) * (dy[:, None] * w3[None, :])
    g[OB2:OW3] = g2.sum(0)
    g2r = r(g2)
    g[OW2:OB2] = (g2r.T @ h1).ravel()
    dh1 = g2r @ r(W2)
    g1 = (z1 > 0) * dh1
    g[OB1:OW2] = g1.sum(0)
    g[OW1:OB1] = (r(g1).T @ xn).ravel()
    return g, float(np.mean((y - t) ** 2))
