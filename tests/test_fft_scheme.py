"""CPU model of the index scheme used by the LDS FFT kernels (phaser_amd/csrc/phx_fft.h):
four-step n = R1*R2 transforms, forward natural -> scrambled, inverse scrambled -> natural,
all stages in place.  The same formulas as the kernel are restated here with numpy so an
indexing mistake fails on the CPU before any GPU run."""
import numpy
import pytest

SHAPES = {16: (4, 4), 32: (8, 4), 64: (8, 8), 128: (16, 8)}


def _dft(v, inv):
    return numpy.fft.ifft(v) * len(v) if inv else numpy.fft.fft(v)


def rows_f1(s, n):
    R1, R2 = SHAPES[n]
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(n) / n)
    for r in range(n):
        for n2 in range(R2):
            idx = R2 * numpy.arange(R1) + n2
            v = _dft(s[r, idx], False)
            v = v * tw[(n2 * numpy.arange(R1))]
            s[r, R2 * numpy.arange(R1) + n2] = v


def rows_f2(s, n, inv=False):
    R1, R2 = SHAPES[n]
    for r in range(n):
        for k1 in range(R1):
            idx = R2 * k1 + numpy.arange(R2)
            s[r, idx] = _dft(s[r, idx], inv)


def rows_iB(s, n):
    R1, R2 = SHAPES[n]
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(n) / n)
    for r in range(n):
        for n2 in range(R2):
            j = numpy.arange(R1)
            v = s[r, R2 * j + n2] * numpy.conj(tw[n2 * j])
            s[r, R2 * j + n2] = _dft(v, True)


def forward(a, n):
    s = a.astype(numpy.complex128).copy()
    rows_f1(s, n); rows_f2(s, n)
    s = s.T.copy(); rows_f1(s, n); rows_f2(s, n); s = s.T.copy()
    return s


def inverse(s, n):
    s = s.copy()
    s = s.T.copy(); rows_f2(s, n, inv=True); rows_iB(s, n); s = s.T.copy()
    rows_f2(s, n, inv=True); rows_iB(s, n)
    return s


def scr_to_freq(p, n):
    R1, R2 = SHAPES[n]
    return p // R2 + R1 * (p % R2)


def freq_to_scr(k, n):
    R1, R2 = SHAPES[n]
    return R2 * (k % R1) + k // R1


@pytest.mark.parametrize('n', [16, 32, 64])
def test_four_step_scrambled_forward_and_natural_inverse(n):
    rng = numpy.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    X = numpy.fft.fft2(a)
    s = forward(a, n)
    p = numpy.arange(n)
    k = scr_to_freq(p, n)
    numpy.testing.assert_allclose(s, X[numpy.ix_(k, k)], atol=1e-9)
    numpy.testing.assert_array_equal(freq_to_scr(k, n), p)
    back = inverse(s, n) / (n * n)
    numpy.testing.assert_allclose(back, a, atol=1e-12)


def test_pointwise_between_transforms_equals_fft_convolution():
    n = 32
    rng = numpy.random.default_rng(5)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    h = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    k = scr_to_freq(numpy.arange(n), n)
    h_scr = h[numpy.ix_(k, k)]   # constant stored in scrambled order
    out = inverse(forward(a, n) * h_scr, n) / (n * n)
    numpy.testing.assert_allclose(out, numpy.fft.ifft2(numpy.fft.fft2(a) * h), atol=1e-12)
