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
FFT_SIZES = (16, 32, 64, 128, 192, 256)


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
    p = numpy.arange(n)
    if n == 256:
        return p   # 256 x 256 planes are transformed in split passes and keep natural order
    L = fft_lanes(n)
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


def hermetian_modes(base_probe, y_orders, x_orders):
    """Hermite-Gauss-like modes y^a x^b P about the probe's intensity centroid, each
    orthogonalised against the earlier ones (with the reference's projection
    sum(prev conj(mode))) and normalised (optics.py:97-137)."""
    real = numpy.float32 if base_probe.dtype == numpy.complex64 else numpy.float64
    y_orders, x_orders = numpy.broadcast_arrays(y_orders, x_orders)
    yy, xx = numpy.indices(base_probe.shape, dtype=real)
    mag = base_probe.real**2 + base_probe.imag**2
    com_y, com_x = (numpy.sum(a * mag) / numpy.sum(mag) for a in (yy, xx))
    yy -= com_y
    xx -= com_x
    modes = numpy.empty((y_orders.size, *base_probe.shape), dtype=base_probe.dtype)
    for i, (yp, xp) in enumerate(zip(y_orders.flat, x_orders.flat)):
        mode = yy**yp * xx**xp * base_probe
        for j in range(i):
            mode -= modes[j] * numpy.sum(modes[j] * numpy.conj(mode))
        mode /= numpy.sqrt(numpy.sum(mode.real**2 + mode.imag**2))
        modes[i] = mode
    return modes.reshape((*y_orders.shape, *base_probe.shape))


def make_hermetian_modes(base_probe, n_modes, base_mode_power=0.7, rel_powers=1.0):
    """`n_modes` probe modes from `base_probe` (optics.py:54-94): the base mode carries
    `base_mode_power`, the others split the rest by `rel_powers`; mode i sits on
    anti-diagonal d = ceil((sqrt(8i + 9) - 3)/2) of the (y order, x order) grid."""
    if not (0.0 <= base_mode_power < 1.0):
        raise ValueError(f"Invalid base_mode_power '{base_mode_power}'. Expected a value between 0 and 1")
    real = numpy.float32 if base_probe.dtype == numpy.complex64 else numpy.float64
    rel = numpy.array(rel_powers, dtype=real).ravel()
    rel = numpy.pad(rel, (0, n_modes - rel.size - 1), mode='edge')[:n_modes - 1]
    powers = numpy.concatenate([[base_mode_power], (1.0 - base_mode_power) * rel])
    i = numpy.arange(n_modes)
    diag = numpy.ceil((numpy.sqrt(8 * i + 9) - 3) / 2).astype(numpy.int_)
    xx = (i - diag * (diag + 1) / 2).astype(numpy.int_)
    yy = diag - xx
    modes = hermetian_modes(base_probe, yy, xx)
    modes *= numpy.array(numpy.sqrt(powers)[:, None, None], dtype=modes.dtype)
    return modes


def resample_slices(obj, old_thicknesses, new_thicknesses):
    """Re-slice an object along z keeping the mean potential (object.py:56-158 with its
    defaults: centres aligned, vacuum padding): log(O) per unit thickness interpolated
    linearly between slice centres, times the new thickness, exponentiated.  Multi to
    single slice is the product; single to multi splits log(O) by thickness fraction."""
    obj = numpy.asarray(obj)
    while obj.ndim < 3:
        obj = obj[None]
    old_t = numpy.asarray(old_thicknesses, dtype=numpy.float64)
    new_t = numpy.asarray(new_thicknesses, dtype=numpy.float64)
    if len(new_t) < 2:
        return obj if obj.shape[0] == 1 else numpy.prod(obj, axis=0, keepdims=True)
    if obj.shape[0] == 1:
        frac = (new_t / numpy.sum(new_t))[:, None, None]
        return numpy.exp((numpy.log(obj) * frac).astype(obj.dtype))
    if obj.shape[0] != len(old_t):
        raise ValueError(f"Expected an object with {len(old_t)} slices, instead got object shape {obj.shape}")
    old_zs = numpy.cumsum(old_t) - old_t / 2.
    new_zs = numpy.cumsum(new_t) - new_t / 2. + (numpy.sum(old_t) - numpy.sum(new_t)) / 2.
    pot = (numpy.log(obj) / old_t[:, None, None]).astype(obj.dtype)
    zero = numpy.zeros_like(pot[0])
    old_zs = numpy.concatenate(([old_zs[0] - old_t[0]], old_zs, [old_zs[-1] + old_t[-1]]))
    pot = numpy.concatenate((zero[None], pot, zero[None]), axis=0)
    real = numpy.float32 if obj.dtype == numpy.complex64 else numpy.float64
    idx = numpy.searchsorted(old_zs, new_zs) - 1
    dz = numpy.diff(old_zs)
    new = numpy.empty((len(new_zs), *pot.shape[1:]), dtype=pot.dtype)
    for i, z in enumerate(new_zs):
        if z <= old_zs[0]:
            new[i] = pot[0]
        elif old_zs[-1] <= z:
            new[i] = pot[-1]
        else:
            k = idx[i]
            t = numpy.array(float((z - old_zs[k]) / dz[k]), dtype=real)
            new[i] = ((1 - t) * pot[k] + t * pot[k + 1]).astype(pot.dtype)
    new *= new_t[:, None, None]
    return numpy.exp(new.astype(obj.dtype))
