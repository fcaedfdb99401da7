"""Pins the oracle's tilted propagators (oracle/optics.py:tilt_propagators, reference
phaser/engines/common/simulation.py:193-222) and their use in the models:
  * a tilt whose lateral walk dz tan(tilt) over a slice is a whole number of pixels makes
    the tilted propagation equal the untilted one shifted by that many pixels (Fourier
    shift theorem; pins the sign, the y/x pairing and each slice's own thickness);
  * the per-position ramp times the band-limited propagator equals the reference's other
    statement of a tilted propagator, fresnel_propagator(..., tilt) (optics.py:190-211);
  * the gradient engine's adjoint with per-position tilts against central finite
    differences (amplitude model)."""
import numpy
import pytest

import oracle.gradient as og
import oracle.num as onum
from oracle.optics import fresnel_propagator, make_propagators, tilt_propagators
from tests.helpers import rel_err
from tests.test_oracle_gradient import _fd_complex, _small


def _grid(n=32, dx=0.25):
    samp = onum.Sampling((n, n), sampling=(dx, dx))
    ky, kx = samp.recip_grid(dtype=numpy.float64)
    return samp, ky, kx


def test_tilted_propagation_is_a_whole_pixel_shift():
    n, dx = 32, 0.25
    samp, ky, kx = _grid(n, dx)
    wl = 0.0197
    thick = numpy.array([20.0, 12.5, 30.0])
    props = make_propagators(ky, kx, samp.k_max, wl, thick, bwlim_frac=None)
    # position 0: no tilt; position 1: walks (+2, -3) px over slice 0, so
    # (+2 * 12.5/20, -3 * 12.5/20) = (1.25, -1.875) px over slice 1: use slice 0 only for
    # the whole-pixel check and position 2 for slice 1's thickness
    my, mx = 2, -3
    t1 = numpy.arctan(numpy.array([my, mx]) * dx / thick[0]) * 1e3
    t2 = numpy.arctan(numpy.array([-1, 4]) * dx / thick[1]) * 1e3
    tilts = numpy.stack([numpy.zeros(2), t1, t2])
    tp = tilt_propagators(ky, kx, thick, props, tilts)
    assert tp.shape == (2, 3, n, n)
    rng = numpy.random.default_rng(3)
    w = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    base0 = onum.ifft2(onum.fft2(w) * props[0])
    base1 = onum.ifft2(onum.fft2(w) * props[1])
    assert rel_err(onum.ifft2(onum.fft2(w) * tp[0, 0]), base0) < 1e-12
    # exp(+2 pi i a k) moves the wave by -a: numpy.roll by (-my, -mx)
    assert rel_err(onum.ifft2(onum.fft2(w) * tp[0, 1]), numpy.roll(base0, (-my, -mx), axis=(0, 1))) < 1e-10
    assert rel_err(onum.ifft2(onum.fft2(w) * tp[1, 2]), numpy.roll(base1, (1, -4), axis=(0, 1))) < 1e-10
    # a planted mistake (y and x tangents swapped) fails the check
    bad = tilt_propagators(ky, kx, thick, props, tilts[:, ::-1])
    assert rel_err(onum.ifft2(onum.fft2(w) * bad[0, 1]), numpy.roll(base0, (-my, -mx), axis=(0, 1))) > 0.1


def test_ramp_matches_fresnel_propagator_with_tilt():
    samp, ky, kx = _grid(32, 0.2)
    wl = 0.0197
    thick = numpy.array([20.0, 20.0])
    tilts = numpy.array([[1.5, -0.7], [-3.0, 2.2]])
    props = make_propagators(ky, kx, samp.k_max, wl, thick)
    tp = tilt_propagators(ky, kx, thick, props, tilts)
    bwlim = props[0] != 0
    for b, (ty, tx) in enumerate(tilts):
        # fresnel_propagator's tilt argument is (x, y)
        ref = fresnel_propagator(ky, kx, wl, thick[0], tilt=(tx, ty)) * bwlim
        assert rel_err(tp[0, b], ref) < 1e-12


def test_tilted_adjoint_matches_finite_differences():
    st, pats, mask, nz, ky, kx, props, group = _small('amplitude')
    rng = numpy.random.default_rng(5)
    st.tilt = rng.uniform(-8.0, 8.0, size=st.scan.shape)
    f = lambda: og.model_loss(st, group, pats, mask, nz, ky, kx, props)  # noqa: E731
    loss, g_obj, g_probe, _g_pos = og.model_grad(st, group, pats, mask, nz, ky, kx, props)
    assert loss == pytest.approx(f(), rel=1e-12)
    untilted = st.tilt
    st.tilt = None
    assert abs(f() - loss) > 1e-6 * abs(loss)     # the tilts change the model
    st.tilt = untilted
    S, H, W = st.object.data.shape
    obj_idx = [(s, int(y), int(x)) for s in range(S) for y, x in rng.integers(H // 2 - 4, H // 2 + 4, size=(3, 2))]
    fd = numpy.array([_fd_complex(f, st.object.data, i) for i in obj_idx])
    assert rel_err(numpy.array([g_obj[i] for i in obj_idx]), fd) < 1e-6
    pr_idx = [(m, int(y), int(x)) for m in range(st.probe.data.shape[0]) for y, x in rng.integers(4, 12, size=(3, 2))]
    fd = numpy.array([_fd_complex(f, st.probe.data, i) for i in pr_idx])
    assert rel_err(numpy.array([g_probe[i] for i in pr_idx]), fd) < 1e-6
