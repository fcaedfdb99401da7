"""Pins of the oracle's constraints and cost regularizers (oracle/regularizers.py,
restating phaser/engines/common/regularizers.py): closed forms, brute force and central
finite differences.  Also the Bluestein tables the object filters upload
(phaser_amd.constraints.bluestein_tables), against numpy's DFT."""
from types import SimpleNamespace as NS

import numpy
import pytest

import oracle.regularizers as oreg
from oracle.num import Sampling


def _state(S=3, H=20, W=22, n=16, seed=0, dtype=numpy.complex128):
    rng = numpy.random.default_rng(seed)
    obj = (rng.normal(1, 0.3, (S, H, W)) + 1j * rng.normal(0, 0.3, (S, H, W))).astype(dtype)
    probe = (rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n))).astype(dtype)
    return NS(object=NS(data=obj, thicknesses=numpy.full(S, 20.0), sampling=NS(sampling=numpy.array([0.2, 0.25]))),
              probe=NS(data=probe, sampling=Sampling((n, n), sampling=(0.2, 0.2))), wavelength=0.0197,
              scan=numpy.zeros((6, 7, 2), numpy.float32))


def test_clamp_caps_modulus_and_keeps_phase():
    st = _state()
    obj0 = st.object.data.copy()
    obj0[0, 0, 0] = 0
    st.object.data = obj0.copy()
    c = oreg.ClampObjectAmplitude(amplitude=1.05)
    c.apply_group(None, st, c.init_state(st))
    out = st.object.data
    m = numpy.abs(obj0) > 0
    numpy.testing.assert_allclose(numpy.abs(out[m]), numpy.minimum(numpy.abs(obj0[m]), 1.05), rtol=1e-12)
    numpy.testing.assert_allclose(numpy.angle(out[m]), numpy.angle(obj0[m]), atol=1e-12)
    assert numpy.isnan(out[0, 0, 0])          # 0/0, as numpy (regularizers.py:40-42)


def _dft2(a):
    """Naive 2-D DFT over the last two axes (brute force)."""
    H, W = a.shape[-2:]
    Fy = numpy.exp(-2j * numpy.pi * numpy.outer(numpy.arange(H), numpy.arange(H)) / H)
    Fx = numpy.exp(-2j * numpy.pi * numpy.outer(numpy.arange(W), numpy.arange(W)) / W)
    return Fy @ a @ Fx.T


def test_limit_probe_support_brute_force():
    st = _state(n=16)
    c = oreg.LimitProbeSupport(max_angle=30.0)
    mask = c.init_state(st)
    kmax = 30e-3 / st.wavelength
    ky, kx = numpy.meshgrid(numpy.fft.fftfreq(16, 0.2), numpy.fft.fftfreq(16, 0.2), indexing='ij')
    numpy.testing.assert_array_equal(mask, ky**2 + kx**2 <= kmax**2)
    assert 0 < mask.sum() < mask.size
    P = st.probe.data.copy()
    c.apply_iter(st, mask)
    # brute force: the centred-real-space convention's shifts cancel for a filter
    want = numpy.fft.ifft2(_dft2(P) * mask)
    numpy.testing.assert_allclose(st.probe.data, want, atol=1e-12)
    # idempotent: a probe inside the support passes unchanged
    Q = st.probe.data.copy()
    c.apply_iter(st, mask)
    numpy.testing.assert_allclose(st.probe.data, Q, atol=1e-12)


def _layers_brute(obj, taps, r, w):
    S = obj.shape[0]
    lg = numpy.log(obj)
    out = numpy.empty_like(obj)
    for s in range(S):
        acc = 0
        for j in range(2 * r + 1):
            acc = acc + taps[j] * lg[min(max(s + j - r, 0), S - 1)]
        out[s] = w * numpy.exp(acc) + (1 - w) * obj[s]
    return out


@pytest.mark.parametrize('S,sigma', [(5, 30.0), (2, 50.0), (9, 15.0)])
def test_layers_brute_force_and_fixed_points(S, sigma):
    st = _state(S=S)
    c = oreg.RegularizeLayers(sigma=sigma, weight=0.7)
    taps, r = c.taps(st.object.thicknesses, numpy.float64)
    assert len(taps) == 2 * r + 1 and abs(taps.sum() - 1) < 1e-12 and numpy.allclose(taps, taps[::-1])
    obj0 = st.object.data.copy()
    c.apply_iter(st, None)
    numpy.testing.assert_allclose(st.object.data, _layers_brute(obj0, taps, r, 0.7), rtol=1e-12)
    # a slice-constant object is a fixed point; weight 0 is the identity
    st.object.data = numpy.broadcast_to(obj0[:1], obj0.shape).copy()
    c.apply_iter(st, None)
    numpy.testing.assert_allclose(st.object.data, numpy.broadcast_to(obj0[:1], obj0.shape), rtol=1e-12)
    st.object.data = obj0.copy()
    oreg.RegularizeLayers(sigma=sigma, weight=0.0).apply_iter(st, None)
    numpy.testing.assert_allclose(st.object.data, obj0, rtol=1e-15)


def test_layers_single_slice_is_untouched():
    st = _state(S=1)
    obj0 = st.object.data.copy()
    oreg.RegularizeLayers().apply_iter(st, None)
    numpy.testing.assert_array_equal(st.object.data, obj0)


def test_object_filters_closed_forms():
    st = _state(S=2, H=21, W=26)
    H, W = 21, 26
    y, x = numpy.meshgrid(numpy.arange(H), numpy.arange(W), indexing='ij')
    # plane waves at integer frequencies (ky, kx) = (3, 5) and (9, 11) cycles per grid
    low = numpy.exp(2j * numpy.pi * (3 * y / H + 5 * x / W))
    high = numpy.exp(2j * numpy.pi * (9 * y / H + 11 * x / W))
    st.object.data = numpy.stack([low, low + high])
    lp = oreg.ObjLowPass(max_freq=0.35)
    mask = lp.init_state(st)
    assert mask[3, 5] and not mask[9, 11]
    lp.apply_iter(st, mask)
    numpy.testing.assert_allclose(st.object.data, numpy.stack([low, low]), atol=1e-12)
    # max_freq beyond the corner frequency keeps everything
    st.object.data = numpy.stack([high, low])
    lp2 = oreg.ObjLowPass(max_freq=0.8)
    lp2.apply_iter(st, lp2.init_state(st))
    numpy.testing.assert_allclose(st.object.data, numpy.stack([high, low]), atol=1e-12)
    # Gaussian: a plane wave is scaled by 1 - w (1 - exp(-2 pi^2 sigma^2 k^2))
    st.object.data = numpy.stack([low, high])
    g = oreg.ObjGaussian(sigma=0.3, weight=0.6)
    filt = g.init_state(st)
    dy, dx = st.object.sampling.sampling
    for plane, (fy, fx) in zip(range(2), ((3, 5), (9, 11))):
        k2 = (fy / (H * dy))**2 + (fx / (W * dx))**2
        want = 1 - 0.6 * (1 - numpy.exp(-2 * numpy.pi**2 * 0.3**2 * k2))
        assert abs(filt[fy, fx] - want) < 1e-6
    g.apply_iter(st, filt)
    numpy.testing.assert_allclose(st.object.data[0], low * filt[3, 5], atol=1e-12)
    numpy.testing.assert_allclose(st.object.data[1], high * filt[9, 11], atol=1e-12)
    st.object.data = numpy.stack([low, high])
    g0 = oreg.ObjGaussian(sigma=0.3, weight=0.0)
    g0.apply_iter(st, g0.init_state(st))
    numpy.testing.assert_allclose(st.object.data, numpy.stack([low, high]), atol=1e-12)


def test_object_filter_brute_force():
    st = _state(S=1, H=9, W=13)
    obj0 = st.object.data.copy()
    g = oreg.ObjGaussian(sigma=0.4, weight=0.8)
    filt = g.init_state(st)
    g.apply_iter(st, filt)
    want = numpy.fft.ifft2(_dft2(obj0[0]) * filt)
    numpy.testing.assert_allclose(st.object.data[0], want, atol=1e-12)


@pytest.mark.parametrize('cls', [oreg.ObjL1, oreg.ObjL2, oreg.ObjTikhonov, oreg.LayersTotalVariation,
                                 oreg.LayersTikhonov])
def test_cost_regularizer_gradients_by_finite_differences(cls):
    st = _state(S=3, H=5, W=6, seed=3)
    reg = cls(cost=0.7)
    group = numpy.zeros((2, 10), numpy.int64)           # 10 of the 42 positions
    g = reg.grad_group(group, st)
    rng = numpy.random.default_rng(5)
    obj0 = st.object.data.copy()
    h = 1e-6
    for _ in range(12):
        idx = tuple(int(rng.integers(0, d)) for d in obj0.shape)
        for unit, part in ((1.0, 'real'), (1j, 'imag')):
            st.object.data = obj0.copy()
            st.object.data[idx] += unit * h
            lp, _ = reg.calc_loss_group(group, st, None)
            st.object.data = obj0.copy()
            st.object.data[idx] -= unit * h
            lm, _ = reg.calc_loss_group(group, st, None)
            fd = (lp - lm) / (2 * h)
            # torch convention 2 dL/dconj(z): dL/dx = Re g, dL/dy = Im g
            want = g[idx].real if part == 'real' else g[idx].imag
            assert abs(fd - want) <= 1e-6 * max(1.0, abs(want)), (cls.__name__, idx, part, fd, want)
    st.object.data = obj0


def test_cost_regularizer_closed_forms():
    st = _state(S=2, H=4, W=5)
    group = numpy.zeros((2, 42), numpy.int64)           # the whole scan: cost scale 1
    st.object.data = numpy.ones((2, 4, 5), numpy.complex128)
    for cls in (oreg.ObjL1, oreg.ObjL2, oreg.ObjTikhonov, oreg.LayersTotalVariation, oreg.LayersTikhonov):
        loss, _ = cls(cost=2.0).calc_loss_group(group, st, None)
        assert loss == 0
    st.object.data = numpy.full((2, 4, 5), 1 + 2j)
    assert oreg.ObjL2(cost=1.0).calc_loss_group(group, st, None)[0] == pytest.approx(40 * 4)
    assert oreg.ObjTikhonov(cost=1.0).calc_loss_group(group, st, None)[0] == 0
    st.object.data[1] = 3
    assert oreg.LayersTikhonov(cost=1.0).calc_loss_group(group, st, None)[0] == pytest.approx(20 * 8)
    # the cost scale is group size / positions
    half = numpy.zeros((2, 21), numpy.int64)
    assert oreg.LayersTikhonov(cost=1.0).calc_loss_group(half, st, None)[0] == pytest.approx(80)


@pytest.mark.parametrize('n', [1, 2, 7, 37, 64, 518, 1030])
def test_bluestein_tables_give_the_dft(n):
    """The host tables of phaser_amd.constraints, used as the kernel uses them (chirp,
    zero-padded FFT_L, product with bhat, inverse FFT_L times L, chirp), give numpy's DFT;
    the inverse by conjugation gives the inverse DFT."""
    from phaser_amd.constraints import bluestein_tables
    chirp, bhat, tw, L = bluestein_tables(n)
    assert L >= 2 * n - 1 and L & (L - 1) == 0 and len(tw) == L // 2
    numpy.testing.assert_allclose(tw, numpy.exp(-2j * numpy.pi * numpy.arange(L // 2) / L), atol=1e-7)
    rng = numpy.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    c, b = chirp.astype(numpy.complex128), bhat.astype(numpy.complex128)

    def dft(v):
        a = numpy.zeros(L, numpy.complex128)
        a[:n] = v * c
        conv = numpy.fft.ifft(numpy.fft.fft(a) * b) * L
        return c * conv[:n]
    want = numpy.fft.fft(x)
    numpy.testing.assert_allclose(dft(x), want, atol=1e-5 * numpy.abs(want).max())
    inv = numpy.conj(dft(numpy.conj(want))) / n
    numpy.testing.assert_allclose(inv, x, atol=1e-5 * numpy.abs(x).max())
