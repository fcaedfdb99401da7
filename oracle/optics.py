"""Probe, shift filter and Fresnel propagator (reference phaser/utils/optics.py,
phaser/engines/common/simulation.py:160-222).  Test infrastructure only."""
import numpy

from .misc import FloatKey
from .num import ifft2, abs2, to_complex_dtype, to_real_dtype


def make_focused_probe(ky, kx, wavelength, aperture, *, defocus=0.):
    """Circular aperture (mrad) with defocus, unit norm, real space (optics.py:30-51)."""
    thetay, thetax = ky * wavelength, kx * wavelength
    theta2 = thetay**2 + thetax**2
    phase = (defocus / (2. * wavelength)) * theta2
    probe = numpy.exp(-2.j * numpy.pi * phase)
    mask = theta2 <= (aperture * 1e-3)**2
    probe *= mask
    probe /= numpy.sqrt(numpy.sum(abs2(probe)))
    return ifft2(probe)


def hermetian_modes(base_probe, y_orders, x_orders):
    """Gram-Schmidt orthonormalised y^a x^b * probe modes (optics.py:97-135)."""
    real_dtype = to_real_dtype(base_probe.dtype)
    y_orders, x_orders = numpy.broadcast_arrays(y_orders, x_orders)
    (yy, xx) = numpy.indices(base_probe.shape, dtype=real_dtype)
    base_probe_mag = abs2(base_probe)
    com_y, com_x = (numpy.sum(a * base_probe_mag) / numpy.sum(base_probe_mag) for a in (yy, xx))
    yy -= com_y
    xx -= com_x
    modes = numpy.empty((y_orders.size, *base_probe.shape), dtype=base_probe.dtype)
    for (i, (y_power, x_power)) in enumerate(zip(y_orders.flat, x_orders.flat)):
        mode = yy**y_power * xx**x_power * base_probe
        for prev_i in range(i):
            mode -= modes[prev_i] * numpy.sum(modes[prev_i] * numpy.conj(mode))
        mode /= numpy.sqrt(numpy.sum(abs2(mode)))
        modes[i] = mode
    return modes.reshape((*y_orders.shape, *base_probe.shape))


def make_hermetian_modes(base_probe, n_modes, base_mode_power=0.7, rel_powers=1.0):
    """Hermite-Gauss modes with power split (optics.py:54-94)."""
    if not (0.0 <= base_mode_power < 1.0):
        raise ValueError(f"Invalid base_mode_power '{base_mode_power}'. Expected a value between 0 and 1")
    rel_powers = numpy.array(rel_powers, dtype=to_real_dtype(base_probe.dtype)).ravel()
    rel_powers = numpy.pad(rel_powers, (0, n_modes - rel_powers.size - 1), mode='edge')[:n_modes - 1]
    powers = numpy.concatenate([[base_mode_power], (1.0 - base_mode_power) * rel_powers])
    i = numpy.arange(n_modes)
    diag_num = numpy.ceil((numpy.sqrt(8 * i + 9) - 3) / 2).astype(numpy.int_)
    xx = (i - diag_num * (diag_num + 1) / 2).astype(numpy.int_)
    yy = diag_num - xx
    modes = hermetian_modes(base_probe, yy, xx)
    modes *= numpy.array(numpy.sqrt(powers)[:, None, None], dtype=modes.dtype)
    return modes


def fourier_shift_filter(ky, kx, shifts):
    """exp(-2 pi i (x kx + y ky)) in complex(ky.dtype) (optics.py:151-170)."""
    dtype = to_complex_dtype(ky.dtype)
    shifts = numpy.array(shifts, dtype=ky.dtype)
    y, x = shifts[..., 0], shifts[..., 1]
    return numpy.exp(numpy.array(-2.j * numpy.pi, dtype=dtype) *
                     (numpy.multiply.outer(x, kx) + numpy.multiply.outer(y, ky)))


def fresnel_propagator(ky, kx, wavelength, delta_z, tilt=(0., 0.)):
    """exp(-i pi dz (lambda k^2 - 2(kx tx + ky ty))) (optics.py:190-211)."""
    tiltx, tilty = numpy.tan(tilt[0] * 1e-3), numpy.tan(tilt[1] * 1e-3)
    k2 = ky**2 + kx**2
    return numpy.exp(-1.j * numpy.pi * delta_z * (wavelength * k2 - 2. * (kx * tiltx + ky * tilty))) \
        .astype(to_complex_dtype(k2.dtype))


def make_propagators(ky, kx, k_max, wavelength, thicknesses, bwlim_frac=2 / 3):
    """Per-slice band-limited propagators, (S-1, Ny, Nx) or None (simulation.py:160-190).

    `ky`, `kx` are the probe recip grid in the simulation dtype; `k_max` is the probe
    sampling's k_max (FP64)."""
    complex_dtype = to_complex_dtype(ky.dtype)
    delta_zs = numpy.asarray(thicknesses)[:-1]
    if len(delta_zs) == 0:
        return None
    unique_zs = set(map(FloatKey, delta_zs))
    if bwlim_frac is not None:
        bwlim = numpy.min(k_max) * bwlim_frac
        k2 = ky**2 + kx**2
        bwlim_mask = k2 <= bwlim**2
    else:
        bwlim_mask = numpy.ones(ky.shape, dtype=numpy.bool_)
    props = {z: fresnel_propagator(ky, kx, wavelength, z).astype(complex_dtype) * bwlim_mask for z in unique_zs}
    return numpy.stack([props[FloatKey(z)] for z in delta_zs], axis=0)


def tilt_propagators(ky, kx, thicknesses, props, tilts):
    """Per-position tilt ramps on the propagators (simulation.py:193-222)."""
    if props is None:
        return None
    if tilts is None:
        return props[:, None, ...]
    dtype = ky.dtype
    complex_dtype = to_complex_dtype(dtype)
    delta_zs = numpy.asarray(thicknesses)[:-1]
    tilt_ramps = numpy.exp(2.j * numpy.pi * numpy.multiply.outer(delta_zs, (
        numpy.multiply.outer(numpy.tan(tilts[..., 0] * 1e-3), ky) +
        numpy.multiply.outer(numpy.tan(tilts[..., 1] * 1e-3), kx)
    )))
    return props[(slice(None), *(None,) * (tilts.ndim - 1), Ellipsis)] * tilt_ramps.astype(complex_dtype)
