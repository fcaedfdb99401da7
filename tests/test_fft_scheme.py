"""CPU model of the index scheme of the FFT kernels (phaser_amd/csrc/phx_fft.h).

L lanes (4, or 8 at n = 128) own one row or column; lane q holds points L j + q.  Per
lane: an E = n/L point DFT over j and a twiddle W_n^(q k); then a DFT_L across the lanes.
With 4 lanes every quad does its DFT_4 as two radix-2 exchanges (partner q^2, then q^1)
with per-lane signs and the lane-3 rotation, as the kernel writes them; with 8 lanes an
in-wave exchange gives lane q the values of all lanes for k1 = 2q, 2q+1 and two DFT-8s
finish.  The
forward leaves each lane's results at positions L k + q (scr_to_freq names the frequency
at each); the inverse undoes it.  Restating the kernel's butterflies here with numpy makes a sign, rotation or index
mistake fail on the CPU before any GPU run."""
import numpy
import pytest

from phaser_amd.optics import fft_lanes, scr_to_freq


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


def _lane8_fwd(lanes, n):
    """8 lanes (n = 128), as the kernel: per lane the DFT-16 and twiddle; the in-wave
    exchange gives lane q the values of every lane for k1 = 2q + s (s = 0, 1); two
    DFT-8s over the source lane leave X[2q + s + 16 r] in register 2r + s."""
    E = n // 8
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(n) / n)
    a = [numpy.fft.fft(lanes[q]) * tw[q * numpy.arange(E)] for q in range(8)]
    out = []
    for q in range(8):
        reg = numpy.empty(E, complex)
        for s in range(2):
            y = numpy.fft.fft([a[t][2 * q + s] for t in range(8)])
            reg[s::2] = y
        out.append(reg)
    return out


def _lane8_inv(lanes, n):
    E = n // 8
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(n) / n)
    z = [[numpy.fft.ifft(lanes[q][s::2]) * 8 for s in range(2)] for q in range(8)]   # z[q][s][q'']
    x = [numpy.array([z[k1 >> 1][k1 & 1][qq] for k1 in range(E)]) for qq in range(8)]
    return [numpy.fft.ifft(x[qq] * numpy.conj(tw[qq * numpy.arange(E)])) * E for qq in range(8)]


def _axis_fwd(v, n):
    """natural vector -> vector in scrambled positions (position L k + q from lane q, k)."""
    L = fft_lanes(n)
    out = numpy.empty(n, complex)
    lanes = (_lane_fwd if L == 4 else _lane8_fwd)([v[q::L] for q in range(L)], n)
    for q in range(L):
        out[q::L] = lanes[q]
    return out


def _axis_inv(v, n):
    L = fft_lanes(n)
    out = numpy.empty(n, complex)
    lanes = (_lane_inv if L == 4 else _lane8_inv)([v[q::L] for q in range(L)], n)
    for q in range(L):
        out[q::L] = lanes[q]
    return out


def forward(a, n):
    """rows (row phase) -> transpose through LDS -> columns (column phase)."""
    s = numpy.stack([_axis_fwd(a[r], n) for r in range(n)])          # s[r, scrambled col]
    return numpy.stack([_axis_fwd(s[:, c], n) for c in range(n)]).T  # [scr row, scr col]


def inverse(s, n):
    t = numpy.stack([_axis_inv(s[:, c], n) for c in range(n)]).T     # columns first
    return numpy.stack([_axis_inv(t[r], n) for r in range(n)])


@pytest.mark.parametrize('n', [16, 32, 64, 128, 192])
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
    # n = 128, 8 lanes, E = 16: register k = 2r + s of lane q holds X[2q + s + 16r]
    f = scr_to_freq(128)
    assert list(f[:8]) == [0, 2, 4, 6, 8, 10, 12, 14] and f[8] == 1 and f[16] == 16 and f[127] == 127


def _dft48(x, inverse=False):
    """The kernel's in-register DFT-48 (phx_fft.h dft48): 48 = 3 x 16, decimation in
    time, DFT-16s of x[3a + b], twiddles W_48^(b k), then a DFT-3 per k written as
    a + s, (a - s/2) -+ i (sqrt(3)/2) d."""
    sg = 1.0 if inverse else -1.0
    y = [numpy.fft.ifft(x[b::3]) * 16 if inverse else numpy.fft.fft(x[b::3]) for b in range(3)]
    k = numpy.arange(16)
    y = [y[0]] + [y[b] * numpy.exp(sg * 2j * numpy.pi * b * k / 48) for b in (1, 2)]
    sm, d = y[1] + y[2], y[1] - y[2]
    t = y[0] - 0.5 * sm
    r = (-sg) * -1j * (numpy.sqrt(3) / 2) * d   # -i h d forward, +i h d inverse
    return numpy.concatenate([y[0] + sm, t + r, t - r])


def test_dft48_split_matches_fft():
    rng = numpy.random.default_rng(48)
    x = rng.normal(size=48) + 1j * rng.normal(size=48)
    numpy.testing.assert_allclose(_dft48(x), numpy.fft.fft(x), atol=1e-12)
    numpy.testing.assert_allclose(_dft48(x, inverse=True), numpy.fft.ifft(x) * 48, atol=1e-12)
    # n = 192: 4 lanes of 48 points; register k of lane q holds X[k + 48 bitrev2(q)]
    f = scr_to_freq(192)
    assert list(f[:4]) == [0, 96, 48, 144] and f[4] == 1 and f[191] == 191


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


def _line256(x, inverse=False):
    """One line of fft256_pass_k (phx_cutout.hip): thread t holds x[16 a + t] and runs a
    DFT_16 over a, the twiddle W_256^(t k1), writes the block slot [k1][t]; thread t then
    plays k1 = t, reads slots [t][j] over j and runs a second DFT_16, whose register k2
    is X[k1 + 16 k2]."""
    sg = 1.0 if inverse else -1.0
    dft16 = (lambda v: numpy.fft.ifft(v) * 16) if inverse else numpy.fft.fft
    tw = numpy.exp(sg * 2j * numpy.pi * numpy.arange(256) / 256)
    blk = numpy.zeros((16, 16), complex)
    for t in range(16):
        y = dft16(x[16 * numpy.arange(16) + t])
        y = y * tw[(t * numpy.arange(16)) & 255]
        blk[:, t] = y                      # slot k1 * 17 + t
    out = numpy.zeros(256, complex)
    for t in range(16):                    # thread t as k1
        z = dft16(blk[t, :])
        out[t + 16 * numpy.arange(16)] = z
    return out


def _fft2_256(a, inverse=False):
    """Row pass then column pass, with the phaser-convention sign and 1/n scale placed
    where the kernel places them (row-pass input for ifft2, column-pass output for fft2)."""
    n = 256
    ky, kx = numpy.meshgrid(numpy.arange(n), numpy.arange(n), indexing='ij')
    sgn = numpy.where((ky + kx) & 1, -1.0, 1.0) / n
    a = a * sgn if inverse else a
    r = numpy.stack([_line256(a[y], inverse) for y in range(n)])
    c = numpy.stack([_line256(r[:, x], inverse) for x in range(n)]).T
    return c if inverse else c * sgn


def test_fft256_split_plane_scheme_matches_phaser_convention():
    import oracle.num as onum
    rng = numpy.random.default_rng(256)
    x = rng.normal(size=256) + 1j * rng.normal(size=256)
    numpy.testing.assert_allclose(_line256(x), numpy.fft.fft(x), atol=1e-10)
    numpy.testing.assert_allclose(_line256(x, True), numpy.fft.ifft(x) * 256, atol=1e-10)
    a = rng.normal(size=(256, 256)) + 1j * rng.normal(size=(256, 256))
    numpy.testing.assert_allclose(_fft2_256(a), onum.fft2(a), atol=1e-10)
    numpy.testing.assert_allclose(_fft2_256(a, True), onum.ifft2(a), atol=1e-10)


def test_split256_group_passes_compose_a_slice_step():
    """The 256^2 group kernels (phx_split256.h) run a forward slice step as three passes
    over the plane: a row pass (transmission, forward row transforms), a column pass
    (forward column transforms, the propagator in natural order, inverse column
    transforms) and a row pass (inverse row transforms, / n^2).  Composed from the line
    model, this is psi_{i+1} = ifft2(fft2(psi O) p); the backward step's column pass with
    conj(p) is its adjoint."""
    n = 256
    rng = numpy.random.default_rng(7)
    psi, obj, prop = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for _ in range(3))

    def cols(a, inverse=False):
        return numpy.stack([_line256(a[:, x], inverse) for x in range(n)]).T

    def rows(a, inverse=False):
        return numpy.stack([_line256(a[y], inverse) for y in range(n)])
    out = rows(cols(cols(rows(psi * obj)) * prop, True), True) / (n * n)
    numpy.testing.assert_allclose(out, numpy.fft.ifft2(numpy.fft.fft2(psi * obj) * prop), atol=1e-9)
    back = rows(cols(cols(rows(out)) * prop.conj(), True), True) / (n * n)
    numpy.testing.assert_allclose(numpy.vdot(back, psi * obj), numpy.vdot(out, out), rtol=1e-9)


def _bfly32(u, v, k):
    """phx_fft.h bfly32 in float32: a = u + W v, b = u - W v, W = W_32^k, with the
    rotation in tangent form (g the larger of |Re W|, |Im W|, r the other over g)."""
    f = numpy.float32
    w = numpy.complex64(numpy.exp(-2j * numpy.pi * k / 32))
    if k in (0, 8, 16, 24):
        t = v * w
        return u + t, u - t
    wx, wy = f(w.real), f(w.imag)
    if abs(wx) >= abs(wy):
        r, g = f(wy / wx), wx
        p = (f(v.real - r * v.imag), f(v.imag + r * v.real))
    else:
        r, g = f(wx / wy), wy
        p = (f(r * v.real - v.imag), f(v.real + r * v.imag))
    a = complex(f(u.real + g * p[0]), f(u.imag + g * p[1]))
    b = complex(f(u.real - g * p[0]), f(u.imag - g * p[1]))
    return a, b


def test_tangent_form_butterflies_match_complex_product():
    rng = numpy.random.default_rng(32)
    for k in range(32):
        for _ in range(20):
            u, v = (complex(*rng.normal(size=2)) for _ in range(2))
            w = numpy.exp(-2j * numpy.pi * k / 32)
            a, b = _bfly32(numpy.complex64(u), numpy.complex64(v), k)
            assert abs(a - (u + w * v)) < 1e-5 * (abs(u) + abs(v) + 1)
            assert abs(b - (u - w * v)) < 1e-5 * (abs(u) + abs(v) + 1)
