"""FP64 reference of the surrogate network (test infrastructure).

Forward and gradient of the 6 -> 64 -> 64 -> 4 ReLU network with sigmoid outputs and the
masked binary cross-entropy loss used by libimc_hip's training kernel.  With bf16=True the
operands that the GPU feeds to its BF16 matrix products (inputs, weights, hidden
activations and deltas) are rounded to bfloat16 exactly as the kernels do; the
arithmetic itself stays FP64.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 6, 64, 4
NPARAM = NIN * NH + NH + NH * NH + NH + NH * NOUT + NOUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NIN * NH].reshape(NIN, NH); o += NIN * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH * NOUT].reshape(NH, NOUT); o += NH * NOUT
    b3 = p[o:o + NOUT]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, X, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    X = q(X)
    H1 = q(np.maximum(X @ q(W1) + b1, 0.0))
    H2 = q(np.maximum(H1 @ q(W2) + b2, 0.0))
    Z = H2 @ q(W3) + b3
    return 1.0 / (1.0 + np.exp(-Z)), (X, H1, H2, Z)


def loss_and_grad(p, X, Y, bf=False):
    """Mean masked BCE (outputs 1-3 only for exiting histories, Y[:,0] == 0) and its
    gradient with respect to the parameter vector."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    n = X.shape[0]
    P, (Xq, H1, H2, Z) = forward(p, X, bf)
    Y = np.asarray(Y, dtype=np.float64)
    mask = np.ones_like(Y)
    mask[:, 1:] = (1.0 - Y[:, 0])[:, None]
    bce = np.maximum(Z, 0) - Z * Y + np.log1p(np.exp(-np.abs(Z)))
    loss = float((bce * mask).sum() / n)
    dZ = (P - Y) * mask / n
    if bf:
        dZ = bf16(dZ)
    gW3 = H2.T @ dZ
    gb3 = dZ.sum(0)
    dH2 = (dZ @ q(W3).T) * (H2 > 0)
    if bf:
        dH2 = bf16(dH2)
    gW2 = H1.T @ dH2
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ q(W2).T) * (H1 > 0)
    if bf:
        dH1 = bf16(dH1)
    gW1 = Xq.T @ dH1
    gb1 = dH1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / NIN), (NIN, NH))
    W2 = rng.normal(0, np.sqrt(2 / NH), (NH, NH))
    W3 = rng.normal(0, np.sqrt(1 / NH), (NH, NOUT))
    return pack(W1, rng.normal(0, 0.1, NH), W2, rng.normal(0, 0.1, NH), W3, rng.normal(0, 0.1, NOUT))
