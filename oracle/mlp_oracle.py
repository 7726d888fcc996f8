"""FP64 numpy oracle of the learned-stop MLP -- TEST INFRASTRUCTURE.

8 -> 64 -> 64 -> 1 with ReLU, ReLU and a sigmoid output, mean binary
cross-entropy.  Parameter layout (floats): W1[64][8] b1[64] W2[64][64] b2[64]
w3[64] b3, then 3 padding floats, mu[8], sd[8] (4820 in all); features enter
as (f - mu) / sd.  The reference has no learned component; this restates
the model so that the GPU's BF16-MFMA gradients can be checked.

bf16=True rounds every matrix-product operand to bfloat16 (nearest even)
before an exact fp64 product, the arithmetic the GPU's MFMA tiles perform up
to fp32 accumulation order.
"""
import numpy as np

N_PARAMS = 4801
N_MODEL = 4820


def split(P):
    P = np.asarray(P, np.float64)
    return (P[0:512].reshape(64, 8), P[512:576], P[576:4672].reshape(64, 64), P[4672:4736], P[4736:4800],
            P[4800])


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def normalise(X, model):
    model = np.asarray(model, np.float32)
    mu, sd = model[4804:4812], model[4812:4820]
    return ((np.asarray(X, np.float32) - mu) / sd).astype(np.float32)


def forward_backward(P, Xn, Y, bf16_operands=False):
    """Mean loss, sigmoid outputs and the gradient (4801 floats) for
    normalised inputs Xn [n, 8] and labels Y [n]."""
    W1, b1, W2, b2, w3, b3 = split(P)
    r = bf16 if bf16_operands else (lambda a: np.asarray(a, np.float64))
    X = np.asarray(Xn, np.float64)
    Y = np.asarray(Y, np.float64)
    n = X.shape[0]
    z1 = r(X) @ r(W1).T + b1
    h1 = np.maximum(z1, 0)
    z2 = r(h1) @ r(W2).T + b2
    h2 = np.maximum(z2, 0)
    o = h2 @ w3 + b3
    p = 1 / (1 + np.exp(-o))
    loss = np.mean(np.maximum(o, 0) - Y * o + np.log1p(np.exp(-np.abs(o))))
    do = (p - Y) / n
    g_w3 = h2.T @ do
    g_b3 = do.sum()
    dz2 = np.outer(do, w3) * (h2 > 0)
    g_W2 = r(dz2).T @ r(h1)
    g_b2 = dz2.sum(0)
    dh1 = r(dz2) @ r(W2)
    dz1 = dh1 * (h1 > 0)
    g_W1 = r(dz1).T @ r(X)
    g_b1 = dz1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, p, g


def loss_only(P, Xn, Y):
    return forward_backward(P, Xn, Y)[0]


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    P = np.zeros(N_MODEL, np.float32)
    P[0:512] = rng.standard_normal(512) * np.sqrt(2 / 8)
    P[576:4672] = rng.standard_normal(4096) * np.sqrt(2 / 64)
    P[4736:4800] = rng.standard_normal(64) * np.sqrt(1 / 64)
    P[4812:4820] = 1.0
    return P
