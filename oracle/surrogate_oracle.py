"""CPU oracle of the trace surrogate's forward and backward passes.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels (include/geosurrogate.h, ionosphere-fdtd_amd/surrogate.py); never by
the product.

The network (no reference counterpart; the reference has no learned model):
    x -> depth x [Linear, SiLU] -> h;  a = h Wa + ba;  z0 = SiLU(h Wz + bz)
    z_s (length L, channels C) -> upsample x2 -> conv k (zero pad k/2) -> SiLU
    shape = conv_k(z_last) (one channel)
with convolutions as im2col products, channels last.

``round_bf16=True`` rounds to BF16 exactly where the GPU path does (inputs,
weights, activations and the output gradients entering a product), with the
products accumulated in FP64; ``round_bf16=False`` is plain FP64 and is what
the finite-difference pinning test checks.  Pinning: tests/
test_oracle_pinning.py::test_surrogate_oracle_gradients_match_finite_differences
(central differences, FP64, every parameter tensor).
"""

from __future__ import annotations

import numpy as np


def round_bf16(x):
    """Round FP64/FP32 values to BF16 (nearest even) and return them as FP64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def im2col_up(x, up, k):
    """x (B, L, C) -> (B * L * up, k * C): nearest upsampling then conv taps."""
    B, L, C = x.shape
    lo = L * up
    u = np.repeat(x, up, axis=1)                      # (B, lo, C)
    half = k // 2
    pad = np.zeros((B, lo + 2 * half, C))
    pad[:, half:half + lo] = u
    cols = np.stack([pad[:, j:j + lo] for j in range(k)], axis=2)  # (B, lo, k, C)
    return cols.reshape(B * lo, k * C)


def col2im_up(dcols, B, L, C, up, k):
    """Adjoint of im2col_up."""
    lo = L * up
    half = k // 2
    d = dcols.reshape(B, lo, k, C)
    dpad = np.zeros((B, lo + 2 * half, C))
    for j in range(k):
        dpad[:, j:j + lo] += d[:, :, j]
    du = dpad[:, half:half + lo]
    return du.reshape(B, L, up, C).sum(axis=2)


def forward(params, x, cfg, round_bf16_points=True):
    r = round_bf16 if round_bf16_points else (lambda v: np.asarray(v, dtype=np.float64))
    W = {k: r(v) if k.startswith("W") else np.asarray(v, dtype=np.float64)
         for k, v in params.items()}
    cache = {"W": W}
    h = r(x)
    acts, pres = [h], []
    for i in range(cfg.depth):
        pre = h @ W[f"W{i}"] + W[f"b{i}"]
        h = r(silu(pre))
        pres.append(pre)
        acts.append(h)
    amp = (h @ W["Wa"] + W["ba"])[:, 0]
    pz = h @ W["Wz"] + W["bz"]
    z = r(silu(pz))
    B = x.shape[0]
    zs, cols_l, cpres = [z], [], []
    length = cfg.seq0
    for s in range(cfg.stages + 1):
        last = s == cfg.stages
        up = 1 if last else 2
        zin = zs[-1].reshape(B, length, cfg.channels[s])
        cols = im2col_up(zin, up, cfg.ksize)
        w, b = ("Wf", "bf") if last else (f"Wc{s}", f"bc{s}")
        pre = cols @ W[w] + W[b]
        cols_l.append(cols)
        length *= up
        if last:
            shape = pre.reshape(B, length)
        else:
            cpres.append(pre)
            zs.append(r(silu(pre)))
    cache.update(acts=acts, pres=pres, pz=pz, zs=zs, cols=cols_l, cpres=cpres, batch=B)
    return shape, amp, cache


def backward(cache, d_shape, d_amp, cfg, round_bf16_points=True):
    r = round_bf16 if round_bf16_points else (lambda v: np.asarray(v, dtype=np.float64))
    W, B = cache["W"], cache["batch"]
    G = {}
    g = np.asarray(d_shape, dtype=np.float64).reshape(-1, 1)
    g16 = r(g)
    length = cfg.length
    for s in range(cfg.stages, -1, -1):
        last = s == cfg.stages
        up = 1 if last else 2
        w, b = ("Wf", "bf") if last else (f"Wc{s}", f"bc{s}")
        cols = cache["cols"][s]
        G[w] = cols.T @ g16
        G[b] = g.sum(axis=0)
        dcols = g16 @ W[w].T
        li = length // up
        dz = col2im_up(dcols, B, li, cfg.channels[s], up, cfg.ksize).reshape(B, -1)
        pre = cache["cpres"][s - 1] if s > 0 else cache["pz"]
        g = dz.reshape(pre.shape) * silu_grad(pre)
        g16 = r(g)
        length = li
    h = cache["acts"][-1]
    G["Wz"] = h.T @ g16
    G["bz"] = g.sum(axis=0)
    dh = g16 @ W["Wz"].T
    da = np.asarray(d_amp, dtype=np.float64).reshape(-1, 1)
    da16 = r(da)
    G["Wa"] = h.T @ da16
    G["ba"] = da.sum(axis=0)
    dh = dh + da16 @ W["Wa"].T
    for i in range(cfg.depth - 1, -1, -1):
        d = dh * silu_grad(cache["pres"][i])
        d16 = r(d)
        G[f"W{i}"] = cache["acts"][i].T @ d16
        G[f"b{i}"] = d.sum(axis=0)
        if i > 0:
            dh = d16 @ W[f"W{i}"].T
    return G


def trace_loss(shape, amp, shape_t, amp_t, amp_weight):
    """Loss of geo_trace_loss and its gradients (FP64)."""
    B, L = shape.shape
    d = shape - shape_t
    e = amp - amp_t
    loss = float(np.sum(d * d) / (B * L) + amp_weight * np.sum(e * e) / B)
    return loss, 2.0 * d / (B * L), 2.0 * amp_weight * e / B
