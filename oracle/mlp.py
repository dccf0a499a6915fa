"""CPU restatement of the learned PCISPH warm start's network -- test
infrastructure only (the tests, the smoke check and the benchmark's CPU leg
use it; the library never does).

The warm start has no counterpart in the reference (its pciInit starts every
step from p = 0, LiquidPhysics/cl_kernels/SPHKernels.cl:661-672); this module
states the network the HIP kernels evaluate, with their precision model:

  forward (per row x, 16 features; parameters in the library's order
  W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3):
    xn = bf16((x - mean) * inv_std)        fp32 subtract then multiply
    y1 = bf16(W1) xn + b1                  exact products, fp32 sums
    h1 = bf16(relu(y1))
    y2 = bf16(W2) h1 + b2
    h2 = bf16(relu(y2))
    y  = sum_k bf16(W3)_k h2_k + b3
  loss over n rows: mean(w d^2), d = y - t / tscale, w = over_weight where
  d > 0, else 1
  backward (straight through the roundings, FP32 master weights):
    dy = 2 w d / n;  g_b3 = sum dy;  g_W3 = sum dy h2
    dy2 = [y2 > 0] bf16(W3) dy;  g_b2 = sum dy2
    g_W2 = bf16(dy2) h1^T
    dy1 = [y1 > 0] (bf16(W2)^T bf16(dy2));  g_b1 = sum dy1
    g_W1 = bf16(dy1) xn^T

Here every sum is taken in float64 over the rounded operands; the kernels sum
in FP32 in the matrix cores' order, so GPU parity is within a tolerance.

Pins (tests/test_mlp_oracle.py): bf16() against torch's bfloat16 conversion
(round to nearest even); the unrounded float64 forward pass against torch's
float64 Linear/ReLU modules on the same packed parameters, and its gradient
against torch autograd and central finite differences; the rounded forward and
gradient equal the unrounded ones when every operand is exactly representable
in bf16.
"""
from __future__ import annotations

import numpy as np

IN = 16
HID = 64
PARAMS = IN * HID + HID + HID * HID + HID + HID + 1


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), np.float32(np.nan), out).astype(np.float32)


def unpack(params):
    p = np.asarray(params, np.float32).reshape(PARAMS)
    o = 0
    W1 = p[o:o + HID * IN].reshape(HID, IN); o += HID * IN
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), np.ravel(b1), np.ravel(W2), np.ravel(b2), np.ravel(w3),
                           np.atleast_1d(b3)]).astype(np.float64)


def _r(a, exact):
    return np.asarray(a, np.float64) if exact else bf16(a).astype(np.float64)


def forward(params, fscale, x, exact: bool = False) -> dict:
    """Network output y [n] and intermediates; exact=True skips every rounding
    and works in float64 throughout."""
    W1, b1, W2, b2, w3, b3 = unpack(params) if not exact else (
        np.asarray(a, np.float64) for a in unpack_f64(params))
    x = np.asarray(x).reshape(-1, IN)
    fs = np.asarray(fscale).reshape(2 * IN)
    if exact:
        xn = (x.astype(np.float64) - fs[:IN].astype(np.float64)) * fs[IN:].astype(np.float64)
    else:
        xn = bf16((x.astype(np.float32) - fs[:IN].astype(np.float32)) * fs[IN:].astype(np.float32)).astype(np.float64)
    y1 = xn @ _r(W1, exact).T + np.asarray(b1, np.float64)
    if not exact:
        y1 = y1.astype(np.float32).astype(np.float64)
    h1 = _r(np.maximum(y1, 0.0).astype(np.float64 if exact else np.float32), exact)
    y2 = h1 @ _r(W2, exact).T + np.asarray(b2, np.float64)
    if not exact:
        y2 = y2.astype(np.float32).astype(np.float64)
    h2 = _r(np.maximum(y2, 0.0).astype(np.float64 if exact else np.float32), exact)
    y = h2 @ _r(w3, exact) + float(b3)
    return {"xn": xn, "y1": y1, "h1": h1, "y2": y2, "h2": h2, "y": y}


def unpack_f64(params):
    p = np.asarray(params, np.float64).reshape(PARAMS)
    o = 0
    W1 = p[o:o + HID * IN].reshape(HID, IN); o += HID * IN
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def loss(params, fscale, tscale, x, t, exact: bool = False, over_weight: float = 1.0) -> float:
    y = forward(params, fscale, x, exact)["y"]
    d = y - np.asarray(t, np.float64) / float(tscale)
    w = np.where(d > 0, over_weight, 1.0)
    return float(np.mean(w * d * d))


def grad(params, fscale, tscale, x, t, exact: bool = False, over_weight: float = 1.0):
    """(gradient [PARAMS] float64, loss) of mean(w (y - t/tscale)^2)."""
    F = forward(params, fscale, x, exact)
    W1, b1, W2, b2, w3, b3 = unpack_f64(params) if exact else unpack(params)
    n = F["y"].shape[0]
    d = F["y"] - np.asarray(t, np.float64).reshape(n) / float(tscale)
    w = np.where(d > 0, over_weight, 1.0)
    dy = 2.0 * w * d / n
    g_b3 = dy.sum()
    g_w3 = dy @ F["h2"]
    dy2 = (F["y2"] > 0) * (_r(w3, exact)[None, :] * dy[:, None])
    g_b2 = dy2.sum(0)
    dy2r = dy2 if exact else bf16(dy2.astype(np.float32)).astype(np.float64)
    g_W2 = dy2r.T @ F["h1"]
    dh1 = dy2r @ _r(W2, exact)
    dy1 = (F["y1"] > 0) * dh1
    if not exact:
        dy1 = dy1.astype(np.float32).astype(np.float64)
    g_b1 = dy1.sum(0)
    dy1r = dy1 if exact else bf16(dy1.astype(np.float32)).astype(np.float64)
    g_W1 = dy1r.T @ F["xn"]
    return pack(g_W1, g_b1, g_W2, g_b2, g_w3, g_b3), float(np.mean(w * d * d))


def he_init(seed: int) -> np.ndarray:
    """He-normal weights and zero biases, FP32."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / IN), (HID, IN))
    W2 = rng.normal(0, np.sqrt(2.0 / HID), (HID, HID))
    w3 = rng.normal(0, np.sqrt(1.0 / HID), HID)
    return pack(W1, np.zeros(HID), W2, np.zeros(HID), w3, 0.0).astype(np.float32)
