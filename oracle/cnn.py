"""CPU restatement (numpy, FP64) of the learned start's network and its training step.

TEST INFRASTRUCTURE ONLY: only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module; the product package
``blaze2d_amd`` never does.

The reference has no learned start (its warm vectors come from the previous
k-point, crates/core/src/eigensolver/subspace_prediction.rs:276-380, and enter
the solver in initialize_block, crates/core/src/eigensolver/initialization.rs:
248-309).  What is restated here is the network the HIP library defines
(include/blaze2d_hip.h, bh_cnn_*): forward pass, projector loss, backward pass
and the Adam step, in FP64, on weights the caller may round to BF16 first
(``bf16``) to match the GPU, which stores weight matrices in BF16 and biases
in FP32.

Pinned (tests/test_cnn_oracle.py) by: the periodic convolution against an
FFT circular cross-correlation; the stride-2 convolution as the stride-1 one
sampled at even positions; the loss's closed form for one vector, its value 0
on any basis of the target span and 1 on an orthogonal block, its invariance
under Y -> Y M; and the whole backward pass against central finite
differences of the forward pass.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

CH = 32
LAYERS = 4
EMB_IN = 4
EMB_OUT = LAYERS * CH
RIDGE = 1e-10


def layout(height: int, width: int, n_vectors: int, pw_half: int) -> List[Tuple[str, tuple]]:
    """(name, shape) of every parameter tensor, in the flat-array order of bh_cnn."""
    ng = (2 * pw_half + 1) ** 2
    out = []
    for l in range(LAYERS):
        cin = 2 if l == 0 else CH
        out += [(f"w{l}", (CH, cin, 3, 3)), (f"b{l}", (CH,))]
    out += [("we1", (CH, EMB_IN)), ("be1", (CH,)), ("we2", (EMB_OUT, CH)), ("be2", (EMB_OUT,))]
    F = CH * (height // 8) * (width // 8)
    out += [("wh", (2 * n_vectors * ng, F)), ("bh", (2 * n_vectors * ng,))]
    return out


def unpack(flat: np.ndarray, lay) -> Dict[str, np.ndarray]:
    p, o = {}, 0
    for name, shape in lay:
        n = int(np.prod(shape))
        p[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    assert o == flat.size, (o, flat.size)
    return p


def pack(p: Dict[str, np.ndarray], lay) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in lay])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (round to nearest, ties to even), as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gpu_weights(p: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """The weights as the GPU uses them: matrices in BF16, biases in FP32."""
    return {k: (bf16(v) if k.startswith("w") else np.asarray(v, np.float32).astype(np.float64)) for k, v in p.items()}


def conv3x3(x: np.ndarray, w: np.ndarray, stride: int) -> np.ndarray:
    """out[b,o,y,x] = sum_{c,dy,dx} w[o,c,dy,dx] x[b,c,(s y+dy-1) mod H,(s x+dx-1) mod W]."""
    out = 0.0
    for dy in range(3):
        for dx in range(3):
            xs = np.roll(x, shift=(1 - dy, 1 - dx), axis=(2, 3))[:, :, ::stride, ::stride]
            out = out + np.einsum("oc,bcyx->boyx", w[:, :, dy, dx], xs)
    return out


def conv3x3_backward(x: np.ndarray, w: np.ndarray, g: np.ndarray, stride: int, need_input: bool = True):
    """Gradients (input, weight) of conv3x3 for the output gradient g."""
    gw = np.zeros_like(w)
    gin = np.zeros_like(x) if need_input else None
    for dy in range(3):
        for dx in range(3):
            xs = np.roll(x, shift=(1 - dy, 1 - dx), axis=(2, 3))[:, :, ::stride, ::stride]
            gw[:, :, dy, dx] = np.einsum("boyx,bcyx->oc", g, xs)
            if need_input:
                up = np.zeros_like(x)
                up[:, :, ::stride, ::stride] = np.einsum("oc,boyx->bcyx", w[:, :, dy, dx], g)
                gin += np.roll(up, shift=(dy - 1, dx - 1), axis=(2, 3))
    return gin, gw


def relu(v):
    return np.maximum(v, 0.0)


def forward(p: Dict[str, np.ndarray], grid: np.ndarray, embed: np.ndarray, n_vectors: int, pw_half: int):
    """grid [B,2,H,W], embed [B,4] -> (Y [B, n_vectors, ng] complex, cache)."""
    ng = (2 * pw_half + 1) ** 2
    B = grid.shape[0]
    x = np.asarray(grid, np.float64)
    e0 = np.asarray(embed, np.float64)
    eh = relu(e0 @ p["we1"].T + p["be1"])
    ev = eh @ p["we2"].T + p["be2"]
    acts = []
    a = x
    for l in range(LAYERS):
        s = 1 if l == 0 else 2
        pre = conv3x3(a, p[f"w{l}"], s) + p[f"b{l}"][None, :, None, None] + ev[:, l * CH:(l + 1) * CH, None, None]
        a = relu(pre)
        acts.append(a)
    flat = a.reshape(B, -1)
    yr = flat @ p["wh"].T + p["bh"]
    y = yr.reshape(B, n_vectors, ng, 2)
    Y = y[..., 0] + 1j * y[..., 1]
    return Y, dict(x=x, e0=e0, eh=eh, acts=acts, flat=flat)


def projector_loss(Y: np.ndarray, U: np.ndarray):
    """Per-sample loss 1 - tr(P_Y P_U)/nb and dL/dY as (d/dRe + i d/dIm), Y, U [nb, ng]."""
    nb = Y.shape[0]
    Yc, Uc = Y.T, U.T  # ng x nb columns
    Gy = Yc.conj().T @ Yc
    Gu = Uc.conj().T @ Uc
    Gy = Gy + np.eye(nb) * RIDGE * np.trace(Gy).real / nb
    Gu = Gu + np.eye(nb) * RIDGE * np.trace(Gu).real / nb
    Giy = np.linalg.inv(Gy)
    Giu = np.linalg.inv(Gu)
    PuY = Uc @ (Giu @ (Uc.conj().T @ Yc))
    T = Giy @ (Yc.conj().T @ PuY)
    loss = 1.0 - np.trace(T).real / nb
    M = (PuY - Yc @ T) @ Giy
    grad = (-2.0 / nb) * M  # d/dRe + i d/dIm
    return loss, grad.T


def loss_and_grads(p: Dict[str, np.ndarray], grid, embed, targets, n_vectors: int, pw_half: int):
    """Mean projector loss over the batch and its gradients w.r.t. every parameter."""
    Y, c = forward(p, grid, embed, n_vectors, pw_half)
    B = Y.shape[0]
    losses = np.zeros(B)
    gY = np.zeros_like(Y)
    for b in range(B):
        losses[b], gY[b] = projector_loss(Y[b], targets[b])
    gY /= B
    g = {}
    gyr = np.stack([gY.real, gY.imag], axis=-1).reshape(B, -1)
    g["wh"] = gyr.T @ c["flat"]
    g["bh"] = gyr.sum(0)
    ga = (gyr @ p["wh"]).reshape(c["acts"][-1].shape)
    gev = np.zeros((B, EMB_OUT))
    for l in range(LAYERS - 1, -1, -1):
        s = 1 if l == 0 else 2
        gpre = ga * (c["acts"][l] > 0)
        g[f"b{l}"] = gpre.sum(axis=(0, 2, 3))
        gev[:, l * CH:(l + 1) * CH] = gpre.sum(axis=(2, 3))
        xin = c["x"] if l == 0 else c["acts"][l - 1]
        gin, g[f"w{l}"] = conv3x3_backward(xin, p[f"w{l}"], gpre, s, need_input=l > 0)
        ga = gin
    g["we2"] = gev.T @ c["eh"]
    g["be2"] = gev.sum(0)
    geh = (gev @ p["we2"]) * (c["eh"] > 0)
    g["we1"] = geh.T @ c["e0"]
    g["be1"] = geh.sum(0)
    return losses.mean(), g, Y


def adam_step(params, grads, m, v, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (step counts from 1); returns (params, m, v)."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return params - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
