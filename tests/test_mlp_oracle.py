"""Pins of the warm-start network's CPU restatement (oracle/mlp.py): the BF16
rounding against torch's conversion, the exact float64 forward pass and
gradient against torch's float64 modules and autograd, the gradient against
finite differences, and the rounded paths against the exact ones where every
operand is representable in BF16."""
import numpy as np
import pytest

from oracle import mlp as M


def test_bf16_matches_torch_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    base = rng.normal(0, 1, 4000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 4000)
    # exact ties (low half 0x8000) with even and odd kept bits, and specials
    ties = (np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x7F7F8000, 0x00008000, 0x00018000],
                     np.uint32).view(np.float32))
    special = np.array([0.0, -0.0, np.inf, -np.inf, 1.0, -1.0, 3.4e38, 1e-45], np.float32)
    x = np.concatenate([base.astype(np.float32), ties, special])
    ours = M.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(M.bf16(np.float32(np.nan)))
    # a tie rounds to the even neighbour: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert M.bf16(np.float32(1 + 2.0 ** -8)) == 1.0
    assert M.bf16(np.float32(1 + 3 * 2.0 ** -8)) == np.float32(1 + 2.0 ** -6)


def _problem(seed, n=37):
    rng = np.random.default_rng(seed)
    p = M.he_init(seed).astype(np.float64)
    p[M.IN * M.HID:M.IN * M.HID + M.HID] = rng.normal(0, 0.1, M.HID)  # b1
    o2 = M.IN * M.HID + M.HID + M.HID * M.HID
    p[o2:o2 + M.HID] = rng.normal(0, 0.1, M.HID)  # b2
    p[-1] = 0.3
    fs = np.concatenate([rng.normal(0, 1, M.IN), rng.uniform(0.5, 2.0, M.IN)])
    x = rng.normal(0, 1.5, (n, M.IN))
    t = rng.normal(0, 2.0, n)
    return p, fs, 1.7, x, t


@pytest.mark.parametrize("ow", [1.0, 3.0])
def test_exact_gradient_matches_finite_differences(ow):
    p, fs, ts, x, t = _problem(3)
    g, l0 = M.grad(p, fs, ts, x, t, exact=True, over_weight=ow)
    assert abs(l0 - M.loss(p, fs, ts, x, t, exact=True, over_weight=ow)) < 1e-12
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(M.PARAMS, 60, replace=False), [M.PARAMS - 1, 0, M.IN * M.HID]])
    eps = 1e-6
    for k in idx:
        pp = p.copy(); pp[k] += eps
        pm = p.copy(); pm[k] -= eps
        fd = (M.loss(pp, fs, ts, x, t, True, ow) - M.loss(pm, fs, ts, x, t, True, ow)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_exact_gradient_catches_a_dropped_relu_mask():
    """A planted mistake in a copy of the backward pass (relu' of layer 2
    dropped) is caught by the finite-difference check."""
    p, fs, ts, x, t = _problem(5)
    F = M.forward(p, fs, x, exact=True)
    W1, b1, W2, b2, w3, b3 = M.unpack_f64(p)
    d = F["y"] - t / ts
    dy = 2 * d / len(t)
    wrong_g_b2 = (w3[None, :] * dy[:, None]).sum(0)  # mask missing
    g, _ = M.grad(p, fs, ts, x, t, exact=True)
    o2 = M.IN * M.HID + M.HID + M.HID * M.HID
    assert (F["y2"] <= 0).any()
    assert not np.allclose(wrong_g_b2, g[o2:o2 + M.HID])


def _representable_problem(seed, n=24):
    """Weights, inputs and normalisation chosen so that every operand and
    intermediate the kernels round is exactly representable in bf16."""
    rng = np.random.default_rng(seed)
    W1 = rng.integers(-2, 3, (M.HID, M.IN)) * 0.25
    W2 = rng.integers(-1, 2, (M.HID, M.HID)) * 0.125
    w3 = rng.integers(-2, 3, M.HID) * 0.5
    b1 = rng.integers(-2, 3, M.HID) * 0.25
    b2 = rng.integers(-2, 3, M.HID) * 0.125
    p = M.pack(W1, b1, W2, b2, w3, 0.5)
    fs = np.concatenate([np.zeros(M.IN), np.ones(M.IN)])
    x = rng.integers(-3, 4, (n, M.IN)).astype(np.float64)
    t = rng.integers(-4, 5, n).astype(np.float64) * 0.5
    return p, fs, 1.0, x, t


def test_rounded_forward_equals_exact_when_representable():
    p, fs, ts, x, t = _representable_problem(7)
    a = M.forward(p, fs, x, exact=False)
    b = M.forward(p, fs, x, exact=True)
    for k in ("xn", "y1", "h1", "y2", "h2", "y"):
        np.testing.assert_array_equal(a[k], b[k])
    assert (b["y1"] < 0).any() and (b["y2"] < 0).any()


def test_rounded_gradient_equals_exact_when_representable():
    p, fs, ts, x, _ = _representable_problem(8, n=16)  # 2/n exact
    # residuals of +-1/2, +-1 keep dy, dy2 and W2^T dy2 on short binary grids
    rng = np.random.default_rng(9)
    W1, b1, W2, b2, w3, b3 = M.unpack_f64(p)
    p = M.pack(W1, b1, np.clip(W2, -0.125, 0.125), b2, np.clip(w3, -0.5, 0.5), b3)
    y = M.forward(p, fs, x, exact=True)["y"]
    t = y - rng.choice([-1.0, -0.5, 0.5, 1.0], y.shape[0])
    ga, la = M.grad(p, fs, ts, x, t, exact=False)
    gb, lb = M.grad(p, fs, ts, x, t, exact=True)
    assert la == lb
    np.testing.assert_array_equal(ga, gb)


def test_rounding_moves_results_by_bf16_amounts():
    """On generic weights the rounded network differs from the exact one by
    about the bf16 unit roundoff (2^-9), not by more."""
    p, fs, ts, x, t = _problem(9, n=500)
    a = M.forward(p, fs, x, exact=False)["y"]
    b = M.forward(p, fs, x, exact=True)["y"]
    rel = np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))
    assert 1e-4 < rel < 2e-2


def _torch_net(torch, p, fs):
    """The network restated with torch modules in float64: Linear(16, 64), ReLU,
    Linear(64, 64), ReLU, Linear(64, 1) after (x - mean) * inv_std, with the
    parameters taken from the packed vector in the library's layout (W1 [64][16],
    b1, W2 [64][64], b2, W3 [64], b3)."""
    W1, b1, W2, b2, w3, b3 = (np.asarray(a, np.float64) for a in M.unpack_f64(p))
    net = torch.nn.Sequential(torch.nn.Linear(M.IN, M.HID), torch.nn.ReLU(), torch.nn.Linear(M.HID, M.HID),
                              torch.nn.ReLU(), torch.nn.Linear(M.HID, 1)).double()
    with torch.no_grad():
        net[0].weight.copy_(torch.from_numpy(W1)); net[0].bias.copy_(torch.from_numpy(b1))
        net[2].weight.copy_(torch.from_numpy(W2)); net[2].bias.copy_(torch.from_numpy(b2))
        net[4].weight.copy_(torch.from_numpy(w3.reshape(1, M.HID)))
        net[4].bias.copy_(torch.from_numpy(np.array([b3])))
    mean = torch.from_numpy(np.asarray(fs[:M.IN], np.float64))
    inv = torch.from_numpy(np.asarray(fs[M.IN:], np.float64))
    return net, mean, inv


def test_exact_forward_matches_torch_float64():
    """The oracle's exact forward pass against torch's float64 modules on the
    same packed parameters (pins the layout and the normalisation)."""
    torch = pytest.importorskip("torch")
    p, fs, ts, x, t = _problem(7, n=200)
    net, mean, inv = _torch_net(torch, p, fs)
    ref = net((torch.from_numpy(x) - mean) * inv).detach().numpy().reshape(-1)
    ours = M.forward(p, fs, x, exact=True)["y"]
    np.testing.assert_allclose(ours, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ow", [1.0, 3.0])
def test_exact_gradient_matches_torch_autograd(ow):
    """The oracle's exact gradient of mean(w (y - t / tscale)^2), w = ow where
    the prediction is above the target, against torch autograd in float64."""
    torch = pytest.importorskip("torch")
    p, fs, ts, x, t = _problem(8, n=150)
    net, mean, inv = _torch_net(torch, p, fs)
    y = net((torch.from_numpy(x) - mean) * inv).reshape(-1)
    d = y - torch.from_numpy(np.asarray(t, np.float64)) / ts
    w = torch.where(d > 0, torch.tensor(ow, dtype=torch.float64), torch.tensor(1.0, dtype=torch.float64))
    loss = torch.mean(w.detach() * d * d)
    loss.backward()
    ref = np.concatenate([net[0].weight.grad.numpy().ravel(), net[0].bias.grad.numpy(),
                          net[2].weight.grad.numpy().ravel(), net[2].bias.grad.numpy(),
                          net[4].weight.grad.numpy().ravel(), net[4].bias.grad.numpy()])
    g, l0 = M.grad(p, fs, ts, x, t, exact=True, over_weight=ow)
    assert abs(l0 - float(loss.detach())) <= 1e-12 * max(1.0, abs(l0))
    np.testing.assert_allclose(g, ref, rtol=1e-10, atol=1e-12)
