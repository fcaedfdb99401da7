"""CPU model of the index scheme of the quad FFT kernels (phaser_amd/csrc/phx_fft.h).

Four lanes (a quad) own one row or column; lane q holds points 4j+q.  Per lane: an
E = n/4 point DFT over j, a twiddle W_n^(q k), then a DFT_4 across the quad in two
radix-2 exchange stages (partner q^2, then q^1) written exactly as the kernel writes
them.  The forward leaves lane q holding X[k + E bitrev2(q)] at position 4k + q; the
inverse undoes it.  Restating the kernel's butterflies here with numpy makes a sign,
rotation or index mistake fail on the CPU before any GPU run."""
import numpy
import pytest

from phaser_amd.optics import scr_to_freq


def _lane_fwd(lanes, n):
    """lanes[q] = E-vector of lane q (points 4j+q) -> lanes after fwd()."""
    E = n // 4
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(n) / n)
    x = [numpy.fft.fft(lanes[q]) * tw[q * numpy.arange(E)] for q in range(4)]
    y = []
    for q in range(4):
        s1 = 1.0 if q < 2 else -1.0          # y = s1 x + partner(q^2)
        yy = s1 * x[q] + x[q ^ 2]
        y.append(yy * -1j if q == 3 else yy)  # lane 3 rotated before the q^1 exchange
    return [(-1.0 if q & 1 else 1.0) * y[q] + y[q ^ 1] for q in range(4)]


def _lane_inv(lanes, n):
    E = n // 4
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(n) / n)
    a = []
    for q in range(4):
        aa = (-1.0 if q & 1 else 1.0) * lanes[q] + lanes[q ^ 1]
        a.append(aa * 1j if q == 3 else aa)
    x = [(1.0 if q < 2 else -1.0) * a[q] + a[q ^ 2] for q in range(4)]
    x = [x[q] * numpy.conj(tw[q * numpy.arange(E)]) for q in range(4)]
    return [numpy.fft.ifft(x[q]) * E for q in range(4)]


def _axis_fwd(v, n):
    """natural vector -> vector in scrambled positions (position 4k+q from lane q, k)."""
    out = numpy.empty(n, complex)
    lanes = _lane_fwd([v[q::4] for q in range(4)], n)
    for q in range(4):
        out[q::4] = lanes[q]
    return out


def _axis_inv(v, n):
    out = numpy.empty(n, complex)
    lanes = _lane_inv([v[q::4] for q in range(4)], n)
    for q in range(4):
        out[q::4] = lanes[q]
    return out


def forward(a, n):
    """rows (row phase) -> transpose through LDS -> columns (column phase)."""
    s = numpy.stack([_axis_fwd(a[r], n) for r in range(n)])          # s[r, scrambled col]
    return numpy.stack([_axis_fwd(s[:, c], n) for c in range(n)]).T  # [scr row, scr col]


def inverse(s, n):
    t = numpy.stack([_axis_inv(s[:, c], n) for c in range(n)]).T     # columns first
    return numpy.stack([_axis_inv(t[r], n) for r in range(n)])


@pytest.mark.parametrize('n', [16, 32, 64, 128])
def test_quad_scheme_scrambled_forward_and_natural_inverse(n):
    rng = numpy.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    X = numpy.fft.fft2(a)
    f = scr_to_freq(n)
    assert sorted(f) == list(range(n))
    s = forward(a, n)
    numpy.testing.assert_allclose(s, X[numpy.ix_(f, f)], atol=1e-9 * n * n)
    numpy.testing.assert_allclose(inverse(s, n) / (n * n), a, atol=1e-12)


def test_scrambled_order_small_case_by_hand():
    # n = 16, E = 4: lane 1 holds X[k + 8] (bitrev2(1) = 2), lane 2 holds X[k + 4]
    f = scr_to_freq(16)
    assert list(f[:8]) == [0, 8, 4, 12, 1, 9, 5, 13]


def test_pointwise_between_transforms_equals_fft_convolution():
    n = 32
    rng = numpy.random.default_rng(5)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    f = scr_to_freq(n)
    out = inverse(forward(a, n) * h[numpy.ix_(f, f)], n) / (n * n)
    numpy.testing.assert_allclose(out, numpy.fft.ifft2(numpy.fft.fft2(a) * h), atol=1e-12)


def test_unsupported_side_raises():
    with pytest.raises(ValueError):
        scr_to_freq(48)
