"""Host-side layout logic, on the CPU: the descriptor's ctypes mirror against the C
header, and the band-limited propagators' zero-column wave mask (which waves of a plane
workgroup skip the column passes of a propagation)."""
import os
import re

import numpy

from phaser_amd import _lib
from phaser_amd.device import zero_column_waves
from phaser_amd.optics import fft_lanes, make_propagators, scr_to_freq
from phaser_amd.sampling import Sampling

HEADER = os.path.join(os.path.dirname(__file__), '..', 'include', 'phaser_mi355x.h')


def test_descriptor_mirror_matches_header():
    src = open(HEADER).read()
    body = src[src.index('typedef struct phx_group_desc {'):src.index('} phx_group_desc;')]
    body = re.sub(r'/\*.*?\*/', '', body.split('{', 1)[1], flags=re.S)
    names = [re.findall(r'\w+', decl)[-1] for decl in body.split(';') if decl.strip()]
    assert names == [f[0] for f in _lib.GroupDesc._fields_]


def _brute_mask(props, n):
    """Wave w of the column phase owns scrambled columns [w c, (w + 1) c), c = 64 / L;
    flag it when all of those columns are zero in every propagator."""
    L = fft_lanes(n)
    c = 64 // L
    f = scr_to_freq(n)
    mask = 0
    for w in range(n * L // 64):
        cols = [f[p] for p in range(w * c, (w + 1) * c)]
        if all(numpy.all(props[:, :, kx] == 0) for kx in cols):
            mask |= 1 << w
    return mask


def test_zero_column_waves_of_band_limited_propagators():
    for n in (32, 64, 128, 192):
        samp = Sampling((n, n), sampling=(0.2, 0.2))
        props = make_propagators(samp, 0.0197, numpy.array([20., 20., 20.]), 2 / 3)
        assert zero_column_waves(props, n) == _brute_mask(props, n)
    # n = 128: the 2/3 band limit zeroes |kx| > 42; waves 6..9 hold kx in [48, 80)
    samp = Sampling((128, 128), sampling=(0.2, 0.2))
    props = make_propagators(samp, 0.0197, numpy.array([20., 20.]), 2 / 3)
    assert zero_column_waves(props, 128) == 0b1111000000
    # no band limit, or a single slice: nothing is skipped
    assert zero_column_waves(make_propagators(samp, 0.0197, numpy.array([20., 20.]), None), 128) == 0
    assert zero_column_waves(None, 128) == 0


def test_split_plane_layout_is_natural_order():
    """256^2 planes run in split row/column passes whose spectra stay in natural order:
    the host's frequency map is the identity there (so patterns, mask, propagators and
    the ky/kx tables upload unpermuted), and the twiddle table is W_256^j."""
    from phaser_amd.optics import twiddles
    f = scr_to_freq(256)
    assert (f == numpy.arange(256)).all()
    tw = twiddles(256)
    numpy.testing.assert_allclose(tw, numpy.exp(-2j * numpy.pi * numpy.arange(256) / 256), atol=1e-7)
    assert tw.dtype == numpy.complex64
    # the other sides keep the lane-scrambled order (a permutation, not the identity)
    for n in (16, 32, 64, 128, 192):
        g = scr_to_freq(n)
        assert sorted(g) == list(range(n)) and not (g == numpy.arange(n)).all()
