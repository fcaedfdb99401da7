"""Pins of the MLP restatement (oracle/mlp.py) used to check the learned-initializer
kernels: finite differences and torch autograd for the gradients, torch's bfloat16
conversion for the rounding, closed forms for Adam, the features and the NRMSE."""
import numpy
import pytest

import oracle.mlp as om


def _net(sizes, seed):
    rng = numpy.random.default_rng(seed)
    Ws = [rng.normal(size=(a, b)) / numpy.sqrt(a) for a, b in zip(sizes[:-1], sizes[1:])]
    bs = [rng.normal(size=b) * 0.1 for b in sizes[1:]]
    return Ws, bs


def test_backward_matches_central_differences():
    sizes = (6, 5, 4, 3)
    Ws, bs = _net(sizes, 0)
    rng = numpy.random.default_rng(1)
    X = rng.normal(size=(7, 6))
    T = rng.normal(size=(7, 3))
    Z, acts = om.forward(X, Ws, bs)
    dWs, dbs = om.backward(Z, T, acts, Ws)
    h = 1e-6
    for l in range(3):
        for idx in [(0, 0), (1, 2), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (om.loss(om.forward(X, Wp, bs)[0], T) - om.loss(om.forward(X, Wm, bs)[0], T)) / (2 * h)
            assert dWs[l][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1] += h
        bm[l][1] -= h
        fd = (om.loss(om.forward(X, Ws, bp)[0], T) - om.loss(om.forward(X, Ws, bm)[0], T)) / (2 * h)
        assert dbs[l][1] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_backward_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    sizes = (8, 6, 6, 4)
    Ws, bs = _net(sizes, 2)
    rng = numpy.random.default_rng(3)
    X = rng.normal(size=(5, 8))
    T = rng.normal(size=(5, 4))
    Z, acts = om.forward(X, Ws, bs)
    dWs, dbs = om.backward(Z, T, acts, Ws)
    tw = [torch.tensor(w, requires_grad=True) for w in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    a = torch.tensor(X)
    for l in range(3):
        a = a @ tw[l] + tb[l]
        if l < 2:
            a = torch.relu(a)
    ((a - torch.tensor(T))**2).mean().backward()
    for l in range(3):
        numpy.testing.assert_allclose(dWs[l], tw[l].grad.numpy(), rtol=1e-12, atol=1e-14)
        numpy.testing.assert_allclose(dbs[l], tb[l].grad.numpy(), rtol=1e-12, atol=1e-14)


def test_round_bf16_ties_to_even_and_matches_torch():
    torch = pytest.importorskip("torch")
    vals = numpy.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, -2.5, 3.14159, 1e-30, 65504.0, 1e30], dtype=numpy.float32)
    r = om.round_bf16(vals)
    assert r[1] == 1.0 and r[2] == 1 + 2**-6
    ref = torch.tensor(vals).to(torch.bfloat16).to(torch.float32).numpy()
    numpy.testing.assert_array_equal(r, ref)
    rng = numpy.random.default_rng(4)
    x = (rng.normal(size=10000) * 10.0**rng.integers(-5, 5, 10000)).astype(numpy.float32)
    numpy.testing.assert_array_equal(om.round_bf16(x), torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy())


def test_adam_first_step_closed_form():
    rng = numpy.random.default_rng(5)
    w = rng.normal(size=20)
    g = rng.normal(size=20)
    w0 = w.copy()
    m = numpy.zeros(20)
    v = numpy.zeros(20)
    om.adam(w, g, m, v, 1, lr=1e-3, eps=1e-8)
    numpy.testing.assert_allclose(w, w0 - 1e-3 * g / (numpy.abs(g) + 1e-8), rtol=1e-12)
    # second step with the same gradient: bias-corrected moments are still g and g^2
    w1 = w.copy()
    om.adam(w, g, m, v, 2, lr=1e-3, eps=1e-8)
    numpy.testing.assert_allclose(w, w1 - 1e-3 * g / (numpy.abs(g) + 1e-8), rtol=1e-10)


def test_features_closed_form():
    p = numpy.arange(16, dtype=numpy.float64).reshape(4, 4)
    f = om.features(p, 0.5)
    bins = numpy.array([[0 + 1 + 4 + 5, 2 + 3 + 6 + 7], [8 + 9 + 12 + 13, 10 + 11 + 14 + 15]])
    numpy.testing.assert_allclose(f, numpy.log1p(bins * 0.5).ravel())
    # 3 x 3 bins of a 6 x 6 pattern: block (i, j) sums 6 (3 i + a) + 3 j + b over a, b < 3
    p6 = numpy.arange(36, dtype=numpy.float64).reshape(6, 6)
    f3 = om.features(p6, 1.0, bin=3)
    blocks = [sum(6 * (3 * i + a) + 3 * j + b for a in range(3) for b in range(3)) for i in range(2) for j in range(2)]
    assert blocks == [63, 90, 225, 252]
    numpy.testing.assert_allclose(f3, numpy.log1p(numpy.array(blocks, dtype=numpy.float64)))


def test_nrmse_phase_invariance_and_value():
    rng = numpy.random.default_rng(6)
    t = numpy.exp(1j * rng.uniform(-0.3, 0.3, (2, 10, 10)))
    assert om.nrmse(t * numpy.exp(0.7j), t) < 1e-12
    t2 = t.copy()
    t2[0] *= 1.1   # the projected object (product over slices) scales by 1.1
    assert om.nrmse(t2, t) == pytest.approx(0.1, rel=1e-12)
    crop = (slice(2, 8), slice(1, 9))
    assert om.nrmse(numpy.ones_like(t), t, crop) > 0


def test_tensor_nrmse_matches_the_oracle():
    """The product-path NRMSE on torch tensors (used where the object is resident on the
    device) equals the oracle's on the same arrays, with and without a crop."""
    import torch
    from phaser_amd.learned import nrmse_tensor
    rng = numpy.random.default_rng(11)
    t = numpy.exp(1j * rng.uniform(-0.3, 0.3, (3, 12, 14))).astype(numpy.complex64)
    o = (t * numpy.exp(1j * rng.normal(0, 0.05, t.shape)) * rng.uniform(0.95, 1.05, t.shape)).astype(numpy.complex64)
    crop = (slice(2, 10), slice(3, 12))
    for c in (None, crop):
        assert nrmse_tensor(torch.from_numpy(o), torch.from_numpy(t), c) == pytest.approx(om.nrmse(o, t, c), rel=1e-12)
    assert nrmse_tensor(torch.from_numpy(t * numpy.exp(0.4j).astype(numpy.complex64)), torch.from_numpy(t)) < 1e-6
