"""Pins of the FP64 network oracle (oracle/surrogate.py): each check is an
independent statement of the mathematics or a library routine, chosen so that a
transposed weight, a wrong tap order, a sign slip or a missing term fails."""
import numpy as np
import pytest

from oracle import surrogate as S


def test_param_count_and_layout():
    assert S.NPARAM == 36357
    p = np.arange(S.NPARAM, dtype=np.float64)
    P = S.unpack(p)
    assert P["W1"].shape == (27, 5, 32) and P["W3"].shape == (27, 32, 5)
    np.testing.assert_array_equal(S.pack(P["W1"], P["b1"], P["W2"], P["b2"], P["W3"], P["b3"]), p)


def test_conv_matches_scipy_correlate():
    ndimage = pytest.importorskip("scipy.ndimage")
    rng = np.random.default_rng(3)
    a = rng.normal(size=(1, 5, 6, 7, 3))
    W = rng.normal(size=(27, 3, 4))
    b = rng.normal(size=4)
    got = S.conv(a, W, b)
    for co in range(4):
        ref = np.full(a.shape[1:4], b[co])
        for ci in range(3):
            ref += ndimage.correlate(a[0, ..., ci], W[:, ci, co].reshape(3, 3, 3), mode="constant", cval=0.0)
        np.testing.assert_allclose(got[0, ..., co], ref, rtol=1e-12, atol=1e-12)


def test_conv_batch_elements_do_not_mix():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(2, 4, 4, 4, 2))
    W = rng.normal(size=(27, 2, 3))
    both = S.conv(a, W)
    np.testing.assert_allclose(both[1], S.conv(a[1:], W)[0], rtol=1e-13, atol=1e-13)


def test_input_gradient_is_the_adjoint():
    rng = np.random.default_rng(5)
    a = rng.normal(size=(2, 4, 5, 6, 3))
    b = rng.normal(size=(2, 4, 5, 6, 7))
    W = rng.normal(size=(27, 3, 7))
    lhs = np.sum(S.conv(a, W) * b)
    rhs = np.sum(a * S.conv_input_grad(b, W))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_weight_gradient_by_linearity():
    rng = np.random.default_rng(6)
    a = rng.normal(size=(1, 4, 5, 6, 3))
    b = rng.normal(size=(1, 4, 5, 6, 7))
    W = rng.normal(size=(27, 3, 7))
    bias = rng.normal(size=7)
    dW, db = S.conv_weight_grad(a, b)
    lhs = np.sum(S.conv(a, W, bias) * b)
    assert abs(lhs - (np.sum(W * dW) + np.sum(bias * db))) <= 1e-11 * abs(lhs)


def test_bf16_matches_torch_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    u = rng.integers(0, 2 ** 32, size=20000, dtype=np.uint64).astype(np.uint32)
    u[:5000] = (u[:5000] & 0xFFFF0000) | 0x8000          # exact ties
    u[5000:8000] = (u[5000:8000] & 0xFFFF0000) | 0x7FFF  # just below a tie
    f = u.view(np.float32)
    f = f[np.isfinite(f) & (np.abs(f) < 1e38)]
    ref = torch.from_numpy(f.copy()).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(S.bf16(f), ref)


def test_gelu_derivative():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (S.gelu(z + h) - S.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(S.gelu_grad(z), fd, rtol=1e-8, atol=1e-9)
    # closed forms: gelu(0) = 0, gelu'(0) = 1/2, gelu(z) -> z for large z
    assert S.gelu(0.0) == 0.0 and abs(S.gelu_grad(0.0) - 0.5) < 1e-15 and abs(S.gelu(10.0) - 10.0) < 1e-12


@pytest.mark.parametrize("gain", [False, True, "scalar"])
def test_gradient_matches_finite_differences(gain):
    """The three losses: y against the target, the modulated g (1 + y) and g (1 + y_0) against it."""
    rng = np.random.default_rng(8)
    p = S.random_params(1, scale=1.5)
    x = rng.normal(scale=0.3, size=(1, 4, 5, 6, 5))
    t = rng.normal(scale=0.1, size=x.shape)
    m = (rng.random(x.shape[:4]) < 0.7).astype(np.float64)
    mod = rng.normal(scale=0.5, size=x.shape) if gain else None
    sc = gain == "scalar"
    loss, g, f = S.loss_and_grad(p, x, t, m, emulate=False, mod=mod, scalar=sc)
    idx = np.concatenate([rng.choice(np.arange(a, b), min(8, b - a), replace=False) for a, b in
                          ((S.OFF_W1, S.OFF_B1), (S.OFF_B1, S.OFF_W2), (S.OFF_W2, S.OFF_B2),
                           (S.OFF_B2, S.OFF_W3), (S.OFF_W3, S.OFF_B3), (S.OFF_B3, S.NPARAM))])
    for i in idx:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (S.loss_and_grad(pp, x, t, m, emulate=False, mod=mod, scalar=sc)[0] -
              S.loss_and_grad(pm, x, t, m, emulate=False, mod=mod, scalar=sc)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * abs(loss)), (i, fd, g[i])


def test_emulated_rounding_only_perturbs():
    rng = np.random.default_rng(9)
    p = S.random_params(2)
    x = rng.normal(scale=0.3, size=(1, 4, 4, 8, 5))
    t = rng.normal(scale=0.1, size=x.shape)
    l0, g0, f0 = S.loss_and_grad(p, x, t, emulate=False)
    l1, g1, f1 = S.loss_and_grad(p, x, t, emulate=True)
    assert abs(l1 - l0) <= 0.05 * l0
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)
    assert not np.array_equal(f0["y"], f1["y"])


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    p = rng.normal(size=50)
    w = torch.tensor(p.copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([w], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in range(1, 8):
        g = rng.normal(size=50) * (0.1 if t % 2 else 2.0)
        w.grad = torch.tensor(g)
        opt.step()
        p, m, v = S.adam(p, g, m, v, t, 3e-3)
        np.testing.assert_allclose(p, w.detach().numpy(), rtol=1e-13, atol=1e-15)
