"""Host-side constants of the forward model: electron wavelength, focused probe, Fresnel
propagators with the band limit, FFT twiddles and the scrambled k-grid tables.

These are computed once per engine with numpy using the reference's expressions and
dtypes (phaser/utils/optics.py:30-51, 190-211; phaser/engines/common/simulation.py:160-190;
phaser/utils/physics.py:30-81), so the propagator constants are bit-identical to the
numpy backend's.  Everything per-group runs in the HIP kernels.
"""
import math

import numpy

_E_REST = 5.1099906e5   # eV
_HC = 1.23984244e4      # eV * angstrom

# plane sides the FFT kernels are built for (phaser_amd/csrc/phx_fft.h)
FFT_SIZES = (16, 32, 64, 128, 192)


def electron_wavelength(energy_ev: float) -> float:
    """hc / sqrt(E (2 m c^2 + E)) in angstrom (physics.py:57-64)."""
    return _HC / math.sqrt(energy_ev * (2 * _E_REST + energy_ev))


def _ifft2_centered(a):
    return numpy.fft.fftshift(numpy.fft.ifft2(a, norm='ortho'), axes=(-2, -1))


def make_focused_probe(ky, kx, wavelength, aperture_mrad, *, defocus=0.):
    """Circular-aperture probe with defocus, unit intensity, real space (optics.py:30-51)."""
    thetay, thetax = ky * wavelength, kx * wavelength
    theta2 = thetay**2 + thetax**2
    phase = (defocus / (2. * wavelength)) * theta2
    probe = numpy.exp(-2.j * numpy.pi * phase)
    probe *= theta2 <= (aperture_mrad * 1e-3)**2
    probe /= numpy.sqrt(numpy.sum(probe.real**2 + probe.imag**2))
    return _ifft2_centered(probe)


def fresnel_propagator(ky, kx, wavelength, delta_z):
    """exp(-i pi dz (lambda k^2 - 2 (kx tan tx + ky tan ty))) in complex(ky.dtype)
    (optics.py:190-211) with zero tilt.  The tilt tangents are numpy float64 scalars, as
    in the reference, so the phase is evaluated in FP64 before the complex64 cast."""
    tiltx, tilty = numpy.tan(0.0 * 1e-3), numpy.tan(0.0 * 1e-3)
    k2 = ky**2 + kx**2
    cdt = numpy.complex64 if k2.dtype == numpy.float32 else numpy.complex128
    return numpy.exp(-1.j * numpy.pi * float(delta_z) * (wavelength * k2 - 2. * (kx * tiltx + ky * tilty))).astype(cdt)


def make_propagators(sampling, wavelength, thicknesses, bwlim_frac=2 / 3, dtype=numpy.float32):
    """(S-1, ny, nx) band-limited propagators, or None for one slice (simulation.py:160-190).

    Unique thicknesses are deduplicated at 5 decimals like FloatKey (misc.py:211-218)."""
    delta_zs = numpy.asarray(thicknesses)[:-1]
    if len(delta_zs) == 0:
        return None
    ky, kx = sampling.recip_grid(dtype=dtype)
    if bwlim_frac is not None:
        bwlim = numpy.min(sampling.k_max) * bwlim_frac
        mask = (ky**2 + kx**2) <= bwlim**2
    else:
        mask = numpy.ones(ky.shape, dtype=numpy.bool_)
    cdt = numpy.complex64 if numpy.dtype(dtype) == numpy.float32 else numpy.complex128
    cache = {}
    out = []
    for z in delta_zs:
        key = round(float(z), 5)
        if key not in cache:
            cache[key] = fresnel_propagator(ky, kx, wavelength, z).astype(cdt) * mask
        out.append(cache[key])
    return numpy.stack(out, axis=0)


def fft_lanes(n: int) -> int:
    """Lanes per row/column of the FFT kernels for an n-point axis (8 at n = 128, else 4)."""
    return 8 if n == 128 else 4


def scr_to_freq(n: int) -> numpy.ndarray:
    """Frequency index held at each scrambled position p = L k + q of an n-point axis
    (register k of lane q after a forward transform).

    4 lanes: X[k + (n/4) bitrev2(q)].  8 lanes (n = 128): register k = 2 r + s of lane q
    holds X[2 q + s + 16 r]."""
    if n not in FFT_SIZES:
        raise ValueError(f"unsupported plane side {n}; the FFT kernels support {FFT_SIZES}")
    L = fft_lanes(n)
    p = numpy.arange(n)
    q, k = p % L, p // L
    if L == 8:
        return 2 * q + (k & 1) + 16 * (k >> 1)
    return k + (n // 4) * (((q & 1) << 1) | (q >> 1))


def scramble(a: numpy.ndarray) -> numpy.ndarray:
    """Reorder the last two (frequency) axes of `a` into the kernels' scrambled order."""
    n = a.shape[-1]
    f = scr_to_freq(n)
    return a[..., f, :][..., :, f]


def twiddles(n: int) -> numpy.ndarray:
    """exp(-2 pi i j / n), j < n, rounded from FP64 to complex64."""
    return numpy.exp(-2j * numpy.pi * numpy.arange(n) / n).astype(numpy.complex64)
