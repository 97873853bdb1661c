"""FP64 restatement of the membrane-substepper MLP (36 -> 128 -> 128 -> 128
-> 24, tanh hidden units) with an analytic backward pass.

Test infrastructure only (imported by tests/).  The network has no
counterpart in the reference; its definition is the one declared in
include/gpucell/surrogate.h.  With `emulate=True` the inputs of every
product (x, h1..h3, the incoming gradient and each layer's backpropagated
gradient) and the weights are rounded to BF16 (round to nearest even), as
the matrix-core kernels do; the arithmetic itself stays FP64, so the GPU
differs only by FP32 accumulation order and its hardware tanh.

Pinned by central differences of the FP64 loss (tests/test_oracle_mlp.py)
and by closed forms of degenerate networks (zero weights: y = b4 and all
weight gradients vanish except those of W4 and b4).
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 36, 128, 24
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H, H)), ("b3", (H,)), ("W4", (OUT, H)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(a) -> np.ndarray:
    """Round to BF16 (nearest even) through the FP32 bit pattern; FP64 out."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unflatten(p) -> dict:
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def flatten(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[n], dtype=np.float64).ravel() for n, _ in SHAPES])


def forward(p, x, emulate: bool = True):
    """y and the cache (inputs of each product) for a batch x [B][36]."""
    w = unflatten(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    # biases are FP32 on the GPU
    bt = np.float32 if emulate else np.float64
    b = {k: np.asarray(w[k], bt).astype(np.float64) for k in ("b1", "b2", "b3", "b4")}
    Wq = {k: r(w[k]) for k in ("W1", "W2", "W3", "W4")}
    x0 = r(np.asarray(x, np.float32) if emulate else x)
    h1 = r(np.tanh(x0 @ Wq["W1"].T + b["b1"]))
    h2 = r(np.tanh(h1 @ Wq["W2"].T + b["b2"]))
    h3 = r(np.tanh(h2 @ Wq["W3"].T + b["b3"]))
    y = h3 @ Wq["W4"].T + b["b4"]
    return y, {"x": x0, "h1": h1, "h2": h2, "h3": h3, "W": Wq}


def backward(p, x, dy, emulate: bool = True) -> np.ndarray:
    """Flat gradient of L = sum(dy * y)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    _, c = forward(p, x, emulate)
    W = c["W"]
    g4 = r(dy)
    gW4, gb4 = g4.T @ c["h3"], g4.sum(axis=0)
    g3 = r((g4 @ W["W4"]) * (1.0 - c["h3"] ** 2))
    gW3, gb3 = g3.T @ c["h2"], g3.sum(axis=0)
    g2 = r((g3 @ W["W3"]) * (1.0 - c["h2"] ** 2))
    gW2, gb2 = g2.T @ c["h1"], g2.sum(axis=0)
    g1 = r((g2 @ W["W2"]) * (1.0 - c["h1"] ** 2))
    gW1, gb1 = g1.T @ c["x"], g1.sum(axis=0)
    return flatten({"W1": gW1, "b1": gb1, "W2": gW2, "b2": gb2, "W3": gW3, "b3": gb3,
                    "W4": gW4, "b4": gb4})


def glorot(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shp[0] + shp[1]))
            d[name] = rng.uniform(-a, a, size=shp)
        else:
            d[name] = np.zeros(shp)
    return flatten(d)
