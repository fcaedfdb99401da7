"""Pins the CPU oracle against the reference's own known-answer tests (transcribed as
data) and against closed forms.  These are the parity anchors of every GPU test."""
import numpy
import pytest
from numpy.testing import assert_array_almost_equal, assert_array_equal

from oracle import misc, num, object as oobj, optics, physics_scan


# The groupings KATs run on the oracle's restatement and on the product's own copy
# (phaser_amd/groups.py), which the engines use for their group order.
def _groups_modules():
    import phaser_amd.groups as pgroups
    return [misc, pgroups]


# --- tests/test_misc.py:7-31 (sparse groupings, seed 'seed1') -----------------------
@pytest.mark.parametrize('misc', _groups_modules(), ids=['oracle', 'product'])
def test_sparse_groupings_known_answer(misc):
    assert [a.tolist() for a in misc.create_sparse_groupings(32, 8, 'seed1')] == [
        [[23, 31, 0, 30, 29, 3, 4, 15]],
        [[14, 6, 9, 28, 13, 18, 17, 26]],
        [[24, 19, 10, 27, 5, 8, 20, 22]],
        [[1, 12, 2, 11, 7, 21, 25, 16]],
    ]
    assert [a.tolist() for a in misc.create_sparse_groupings(33, 8, 'seed1')] == [
        [[24, 16, 0, 30, 25, 3, 4]],
        [[15, 14, 6, 9, 7, 13, 21]],
        [[19, 31, 28, 17, 23, 18, 10]],
        [[29, 27, 32, 5, 8, 20]],
        [[22, 1, 26, 12, 2, 11]],
    ]
    arr = numpy.zeros((32, 32, 4))
    g = misc.create_sparse_groupings(arr.shape, 128)
    assert g[0].shape == (3, 128)
    assert arr[tuple(g[0])].shape == (128,)


# --- tests/test_misc.py:33-49 (compact groupings) -----------------------------------
@pytest.mark.parametrize('misc', _groups_modules(), ids=['oracle', 'product'])
def test_compact_groupings_known_answer(misc):
    scan = numpy.stack(numpy.meshgrid(numpy.linspace(0., 20., 8), numpy.linspace(0., 20., 8),
                                      indexing='ij'), axis=-1).reshape(-1, 2)
    assert [a.tolist() for a in misc.create_compact_groupings(scan, 16, 'seed1')] == [
        [[0, 1, 2, 3, 8, 9, 10, 11, 16, 17, 18, 19, 24, 25, 26, 39]],
        [[4, 5, 6, 7, 12, 13, 14, 15, 20, 21, 22, 23, 27, 28, 29, 30]],
        [[32, 33, 34, 40, 41, 42, 43, 48, 49, 50, 51, 56, 57, 58, 59, 63]],
        [[31, 35, 36, 37, 38, 44, 45, 46, 47, 52, 53, 54, 55, 60, 61, 62]],
    ]


@pytest.mark.parametrize('misc', _groups_modules(), ids=['oracle', 'product'])
def test_mask_fraction_of_groups(misc):
    # every floor(n/k)-th group starting at 0, k = ceil(n * fraction)
    assert misc.mask_fraction_of_groups(64, 0.1).nonzero()[0].tolist() == [0, 9, 18, 27, 36, 45, 54, 63]
    assert misc.mask_fraction_of_groups(4, 0.1).nonzero()[0].tolist() == [0]
    assert misc.mask_fraction_of_groups(3, 1.0).all()


# --- tests/test_object.py:12-29 (random_phase_object, seed 2620771887) -------------
RPO = numpy.array([
    [1.-1.5272086e-05j, 1.+1.0225522e-04j, 1.-8.0865902e-05j, 1.-1.7328106e-05j, 1.-1.2898073e-04j, 1.+2.2908196e-05j, 1.+8.1173976e-06j, 1.+2.1377344e-05j],
    [1.+7.4363430e-05j, 1.-9.1323782e-05j, 1.-2.0272582e-04j, 1.-4.8823396e-05j, 1.+9.3021641e-05j, 1.+1.0718761e-04j, 1.+5.0221975e-06j, 1.-5.5743083e-05j],
    [1.+9.5888179e-05j, 1.+7.0838556e-05j, 1.-1.1567964e-04j, 1.+1.3202346e-04j, 1.+1.3625837e-04j, 1.-5.2489726e-05j, 1.-1.3756646e-04j, 1.+2.8579381e-05j],
    [1.-6.3651263e-05j, 1.-4.5127890e-05j, 1.+5.5954431e-05j, 1.-1.8197308e-04j, 1.+6.3579530e-05j, 1.-4.6506138e-05j, 1.+5.1510222e-05j, 1.+1.0700211e-04j],
    [0.99999994-2.6713090e-04j, 1.+1.8953861e-04j, 1.+1.1097628e-04j, 1.-5.9648257e-05j, 1.+4.2086729e-05j, 1.+6.0222395e-05j, 1.-8.4926840e-05j, 0.99999994-2.6520275e-04j],
    [1.-4.1160638e-05j, 1.+8.4538617e-05j, 1.+4.1620955e-05j, 1.+1.6012797e-05j, 1.-1.3888512e-05j, 1.+9.1871625e-06j, 1.+5.0595980e-05j, 1.+2.3048995e-04j],
    [1.-2.8506602e-05j, 1.+2.4769653e-05j, 1.-2.3920753e-05j, 1.+8.0796681e-06j, 0.99999994-2.5373933e-04j, 1.+6.9838488e-06j, 1.+3.8624425e-05j, 1.+1.1229565e-04j],
    [1.-4.8519720e-05j, 1.+9.4494520e-05j, 1.+4.9148810e-05j, 1.-1.3229759e-04j, 1.-2.6898948e-05j, 1.-9.8376579e-05j, 1.+6.9485272e-05j, 1.-9.8597156e-05j],
], dtype=numpy.complex64)


def test_random_phase_object_known_answer():
    obj = oobj.random_phase_object((8, 8), 1e-4, seed=2620771887, dtype=numpy.complex64)
    assert obj.dtype == numpy.complex64
    assert_array_almost_equal(obj, RPO)


# --- tests/test_num.py:102-155 (FFT convention closed forms) -----------------------
def test_fft2_closed_forms():
    a = numpy.pad(numpy.array([[5.]], dtype=numpy.float32), ((2, 2), (2, 2)))
    assert_array_almost_equal(num.fft2(a), numpy.full((5, 5), 1., dtype=numpy.complex64))
    a = numpy.full((5, 5), 1 + 1.j, dtype=numpy.complex64)
    assert_array_almost_equal(num.fft2(a), numpy.pad([[5. + 5.j]], ((0, 4), (0, 4))).astype(numpy.complex64))


def test_ifft2_closed_forms():
    a = numpy.pad(numpy.array([[5.]], dtype=numpy.float32), ((0, 4), (0, 4)))
    assert_array_almost_equal(num.ifft2(a), numpy.full((5, 5), 1., dtype=numpy.complex64), decimal=5)
    a = numpy.full((5, 5), 1 + 1.j, dtype=numpy.complex64)
    assert_array_almost_equal(num.ifft2(a), numpy.pad([[5. + 5.j]], ((2, 2), (2, 2))).astype(numpy.complex64), decimal=5)


@pytest.mark.parametrize('n', [8, 6])
def test_fft2_against_naive_dft(n):
    """Brute-force DFT with the centred-real-space convention; catches a dropped shift,
    a wrong sign or the wrong normalisation."""
    rng = numpy.random.default_rng(1)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    c = n // 2  # real-space centre index (ifftshift moves it to 0)
    ys = (numpy.arange(n) - c) % n
    k = numpy.arange(n)
    W = numpy.exp(-2j * numpy.pi * numpy.outer(k, ys) / n)
    expected = W @ a @ W.T / n
    assert_array_almost_equal(num.fft2(a), expected, decimal=10)
    assert_array_almost_equal(num.ifft2(num.fft2(a)), a, decimal=10)
    assert numpy.sum(num.abs2(num.fft2(a))) == pytest.approx(numpy.sum(num.abs2(a)))


def test_even_fft2_is_checkerboard_modulated_fft():
    """For even N, fft2 = (-1)^(ky+kx) * FFT and ifft2 = IFFT((-1)^(ky+kx) * .): the form
    the GPU kernels use."""
    n = 16
    rng = numpy.random.default_rng(2)
    a = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(numpy.complex128)
    ky, kx = numpy.indices((n, n))
    c = (-1.0) ** (ky + kx)
    assert_array_almost_equal(num.fft2(a), c * numpy.fft.fft2(a, norm='ortho'), decimal=12)
    assert_array_almost_equal(num.ifft2(a), numpy.fft.ifft2(c * a, norm='ortho'), decimal=12)


# --- tests/test_object.py:31-144 (ObjectSampling) ------------------------------------
def test_object_sampling_grid():
    samp = oobj.ObjectSampling((16, 15), (0.5, 2.0))
    assert_array_equal(samp.shape, [16, 15])
    assert_array_almost_equal(samp.corner, [-3.75, -14.0])
    assert_array_almost_equal(samp.max, [3.75, 14.0])
    assert samp.mpl_extent() == pytest.approx((-15.0, 15.0, 4.0, -4.0))
    assert samp.mpl_extent(False) == pytest.approx((-14.0, 16.0, 4.25, -3.75))
    yy, xx = samp.grid()
    assert_array_almost_equal(xx[0, :], [-14., -12., -10., -8., -6., -4., -2., 0., 2., 4., 6., 8., 10., 12., 14.])
    assert_array_almost_equal(yy[:, 0], numpy.arange(-3.75, 3.76, 0.5))
    assert samp.get_region_crop() == (slice(None, None), slice(None, None))


def test_object_from_scan():
    scan = numpy.array([[-4.8, -1.], [2.1, 1.], [4.19, 9.9], [0., numpy.nan]])
    samp = oobj.ObjectSampling.from_scan(scan, sampling=[0.5, 0.5], pad=2.)
    assert_array_almost_equal(samp.min, [-6.8, -3.])
    assert_array_almost_equal(samp.max, [6.2, 12.0])
    assert_array_almost_equal(samp.shape, (27, 31))
    assert_array_almost_equal(samp.region_min, [-4.8, -1.])
    assert_array_almost_equal(samp.region_max, [4.19, 9.9])
    assert samp.get_region_crop() == (slice(4, 22), slice(4, 26))


def test_object_expand_to_scan():
    samp = oobj.ObjectSampling(shape=(10, 10), sampling=(10.0, 20.0), corner=(-5.0, -10.0),
                               region_min=None, region_max=(20.0, 500.0))
    scan = numpy.array([[5.0, -10.0], [75.0, 50.0]])
    new = samp.expand_to_scan(scan, pad=15.0)
    assert_array_equal(new.shape, [12, 11])
    assert_array_almost_equal(new.min, [-15.0, -30.0])
    assert_array_almost_equal(new.max, [95.0, 170.0])
    assert new.region_min is None
    assert_array_almost_equal(new.region_max, [75.0, 500.0])


def test_object_slicing():
    samp = oobj.ObjectSampling((40, 31), (0.5, 2.0))
    yy, xx = samp.grid()
    yp, xp = numpy.round(samp._pos_to_object_idx(numpy.array((-1.25, 12.0)), (1, 1))).astype(numpy.int_)
    assert yy[yp, xp] == pytest.approx(-1.25)
    assert xx[yp, xp] == pytest.approx(12.0)
    assert_array_equal(numpy.round(samp._pos_to_object_idx(numpy.array((-1.25 + 0.24, 12.0 - 0.99)), (1, 1))).astype(numpy.int_), (yp, xp))
    n = 2
    pos = numpy.array((-1.25 - 0.24 + 0.5 * n, 12.0 - 0.99 + 2.0 * n))
    shape = (1 + 2 * n, 1 + 2 * n)
    assert_array_equal(numpy.round(samp._pos_to_object_idx(pos, shape)).astype(numpy.int_), (yp, xp))
    assert samp.slice_at_pos(pos, shape) == (slice(yp, yp + 1 + 2 * n), slice(xp, xp + 1 + 2 * n))
    n = 6
    pos = numpy.array((-1.25 - 0.24 + 0.5 * (n - 1 / 2.), 12.0 - 0.99 + 2.0 * (n - 1 / 2.)))
    shape = (2 * n, 2 * n)
    assert_array_equal(numpy.round(samp._pos_to_object_idx(pos, shape)).astype(numpy.int_), (yp, xp))
    assert samp.slice_at_pos(pos, shape) == (slice(yp, yp + 2 * n), slice(xp, xp + 2 * n))


def test_start_index_ties_round_half_to_even():
    """Exact .5 ties round to even (numpy.round), not away from zero."""
    samp = oobj.ObjectSampling((100, 100), (1.0, 1.0), corner=(0.0, 0.0))
    # cutout 64 -> shift -31.5; pos 40 -> idx 8.5 -> 8 ; pos 41 -> 9.5 -> 10
    idx = samp.start_idxs(numpy.array([[40., 40.], [41., 41.]]), (64, 64))
    assert idx.tolist() == [[8, 8], [10, 10]]
    sh = samp.get_subpx_shifts(numpy.array([[40., 40.], [41., 41.]]), (64, 64))
    assert sh.tolist() == [[0.5, 0.5], [-0.5, -0.5]]


# --- tests/test_object.py:266-296 (multi-dim cutout ramp) ----------------------------
def test_cutout_multidim_ramp():
    samp = oobj.ObjectSampling((200, 200), (1.0, 1.0))
    zz, yy, xx = numpy.indices((3, 200, 200), dtype=numpy.float32)
    obj = 1000. * zz + (2. * yy + xx)
    pos = numpy.array([[[-50., -50.], [-50., 50.]], [[50., -50.], [50., 50.]]])
    arr = samp.get_view_at_pos(obj, pos, (80, 100))
    assert arr.shape == (2, 2, 3, 80, 100)
    assert_array_almost_equal(arr[..., 0, 0], [
        [[20., 1020., 2020.], [120., 1120., 2120.]],
        [[220., 1220., 2220.], [320., 1320., 2320.]],
    ])
    assert_array_almost_equal(*numpy.broadcast_arrays(numpy.diff(arr, axis=-1), 1.))
    assert_array_almost_equal(*numpy.broadcast_arrays(numpy.diff(arr, axis=-2), 2.))


ADD_POS = numpy.array([
    [50.60303087, 65.79992213], [55.29351178, -23.20535679], [-24.55795797, 61.03108143],
    [31.11937136, -15.76195155], [28.01925522, 45.1035061], [10.36488598, 2.55497549],
    [-57.42025, 48.55254953], [4.08776065, 50.33652909], [-15.73631518, -36.91599871],
    [-25.16043481, 28.76977663], [31.42525951, 10.31409966], [-19.79372857, 55.79668508],
    [-60.01548957, 26.95982136], [-31.2433725, -54.82789204], [59.78674718, 40.56994336],
    [-46.26502396, -38.71300405], [-66.83498502, -6.33424509], [62.7652186, -12.48322478],
    [-2.31684356, -6.0558], [-32.38632762, 5.55021944], [-52.72306553, -19.39256228],
    [13.48322753, 60.19354572], [-30.44280284, 52.91061857], [-43.27221074, 51.47703124],
    [-24.6619304, 66.08921008], [59.68745301, 43.99038992], [-60.82051963, -42.56576538],
    [-38.95303419, -38.40890994], [-55.76325197, -55.0998433], [48.4540186, 52.91460643],
])


# --- tests/test_object.py:181-244 (add views: sum invariant) -------------------------
def test_add_view_sum_invariant():
    samp = oobj.ObjectSampling((200, 200), (1.0, 1.0))
    obj = numpy.zeros((200, 200), dtype=numpy.float32)
    cut = numpy.full((30, 64, 64), 10., dtype=numpy.float32)
    out = samp.add_view_at_pos(obj, ADD_POS, cut)
    assert numpy.sum(out) == 10. * numpy.prod(cut.shape)
    # brute force: integer counts of coverage
    cnt = numpy.zeros((200, 200), dtype=numpy.int64)
    for p in ADD_POS:
        sy, sx = samp.slice_at_pos(p, (64, 64))
        cnt[sy, sx] += 1
    assert_array_equal(out, 10. * cnt)


# --- optics -------------------------------------------------------------------------
def test_focused_probe_unit_norm_and_propagator_sign():
    """tests/test_optics.py:44-57: -200 A defocus then +200 A propagation = focused."""
    from oracle.num import Sampling
    s = Sampling((256, 256), extent=(25., 25.))
    ky, kx = s.recip_grid(dtype=numpy.float64)
    p_def = optics.make_focused_probe(ky, kx, 0.0251, 10., defocus=-200.)
    assert numpy.sum(num.abs2(p_def)) == pytest.approx(1.0)
    prop = optics.fresnel_propagator(ky, kx, 0.0251, 200.)
    p = num.ifft2(num.fft2(p_def) * prop)
    p_focus = optics.make_focused_probe(ky, kx, 0.0251, 10., defocus=0.)
    assert_array_almost_equal(numpy.abs(p), numpy.abs(p_focus), decimal=10)
    # the opposite sign does not refocus
    p_bad = num.ifft2(num.fft2(p_def) * optics.fresnel_propagator(ky, kx, 0.0251, -200.))
    assert numpy.max(numpy.abs(numpy.abs(p_bad) - numpy.abs(p_focus))) > 1e-3


def test_propagators_band_limit_and_roundtrip():
    from oracle.num import Sampling
    s = Sampling((64, 64), extent=(12.8, 12.8))
    ky, kx = s.recip_grid(dtype=numpy.float32)
    props = optics.make_propagators(ky, kx, s.k_max, 0.0197, numpy.array([20., 20., 20.], dtype=numpy.float32))
    assert props.shape == (2, 64, 64) and props.dtype == numpy.complex64
    k2 = ky.astype(numpy.float64)**2 + kx**2
    inside = k2 <= (numpy.min(s.k_max) * 2 / 3)**2
    assert numpy.all(props[:, ~inside] == 0)
    assert_array_almost_equal(numpy.abs(props[:, inside]), 1.0, decimal=6)
    # single slice -> None (no band limit anywhere)
    assert optics.make_propagators(ky, kx, s.k_max, 0.0197, numpy.array([], dtype=numpy.float32)) is None
    # closed form at one pixel
    iy, ix = 3, 5
    expect = numpy.exp(-1j * numpy.pi * 20. * 0.0197 * (float(ky[iy, ix])**2 + float(kx[iy, ix])**2))
    assert props[0, iy, ix] == pytest.approx(expect, rel=1e-5)


def test_fourier_shift_filter_shifts_by_integer_pixels():
    from oracle.num import Sampling
    s = Sampling((32, 32), sampling=(1.0, 1.0))
    ky, kx = s.recip_grid(dtype=numpy.float64)
    rng = numpy.random.default_rng(3)
    a = rng.normal(size=(32, 32)) + 0j
    f = optics.fourier_shift_filter(ky, kx, numpy.array([[2.0, -3.0]]))[0]
    shifted = num.ifft2(num.fft2(a) * f)
    assert_array_almost_equal(shifted, numpy.roll(a, (2, -3), axis=(0, 1)), decimal=10)


# --- tests/test_physics.py:7-49 ----------------------------------------------------
def test_electron_200kev():
    e = physics_scan.Electron(200e3)
    for k, v in {'total_energy': 710999.06, 'mass': 1.267471e-30, 'momentum': 494367.9034888,
                 'wavelength': 0.02507934741, 'gamma': 1.391390152459, 'beta': 0.69531442627,
                 'velocity': 208450020.93, 'interaction_param': 7.2883988338e-4}.items():
        assert pytest.approx(v, rel=1e-10) == getattr(e, k)


def test_electron_0kev():
    e = physics_scan.Electron(0.)
    assert e.momentum == 0. and e.gamma == 1. and e.beta == 0.
    with pytest.raises(ZeroDivisionError):
        e.wavelength
