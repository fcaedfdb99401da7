"""FFT convention, abs2, Sampling (reference phaser/utils/num.py).  Test infrastructure only."""
import numpy


def fft2(a):
    """Forward FFT over the last two axes, real space centred, ortho norm.

    Follows phaser/utils/num.py:456-468: fft2(a) = FFT(ifftshift(a), norm='ortho').
    """
    return numpy.fft.fft2(numpy.fft.ifftshift(a, axes=(-2, -1)), norm='ortho')


def ifft2(a):
    """Inverse FFT over the last two axes (phaser/utils/num.py:434-441):
    ifft2(a) = fftshift(IFFT(a, norm='ortho'))."""
    return numpy.fft.fftshift(numpy.fft.ifft2(a, norm='ortho'), axes=(-2, -1))


def abs2(x):
    """|x|^2 as x.real**2 + x.imag**2 (phaser/utils/num.py:505-512)."""
    x = numpy.asarray(x)
    return x.real**2. + x.imag**2.


def check_finite(*arrs, context=None):
    """Raise ValueError on NaN/inf (phaser/utils/num.py:535-541)."""
    if not all(numpy.all(numpy.isfinite(arr)) for arr in arrs):
        if context:
            raise ValueError(f"NaN or inf encountered in {context}")
        raise ValueError("NaN or inf encountered")


def to_complex_dtype(dtype):
    return numpy.complex64 if numpy.dtype(dtype) == numpy.float32 else numpy.complex128


def to_real_dtype(dtype):
    d = numpy.dtype(dtype)
    if d in (numpy.complex64, numpy.float32):
        return numpy.float32
    return numpy.float64


class Sampling:
    """Probe/pattern sampling grid (phaser/utils/num.py:544-670)."""

    def __init__(self, shape, *, extent=None, sampling=None):
        self.shape = numpy.broadcast_to(numpy.asarray(shape).astype(numpy.int_), (2,))
        if extent is not None:
            self.extent = numpy.broadcast_to(numpy.asarray(extent).astype(numpy.float64), (2,))
            self.sampling = self.extent / self.shape
        elif sampling is not None:
            self.sampling = numpy.broadcast_to(numpy.asarray(sampling).astype(numpy.float64), (2,))
            self.extent = self.sampling * self.shape
        else:
            raise ValueError("Either 'extent' or 'sampling' must be specified")

    @property
    def k_max(self):
        # num.py:561-566
        return (1 / (2 * self.sampling)).astype(numpy.float64)

    @property
    def corner(self):
        return ((-self.extent + self.sampling) / 2.).astype(numpy.float64)

    def real_grid(self, *, dtype=numpy.float64):
        # num.py:615-627
        corner = self.corner
        ys = numpy.linspace(corner[0], corner[0] + self.extent[0], self.shape[0], endpoint=False, dtype=dtype)
        xs = numpy.linspace(corner[1], corner[1] + self.extent[1], self.shape[1], endpoint=False, dtype=dtype)
        return tuple(numpy.meshgrid(ys, xs, indexing='ij'))

    def recip_grid(self, *, centered=False, dtype=numpy.float64):
        """fftfreq grid cast to `dtype`, zero frequency in the corner (num.py:651-670)."""
        ky = numpy.fft.fftfreq(int(self.shape[0]), float(self.sampling[0])).astype(dtype)
        kx = numpy.fft.fftfreq(int(self.shape[1]), float(self.sampling[1])).astype(dtype)
        if centered:
            ky = numpy.fft.fftshift(ky)
            kx = numpy.fft.fftshift(kx)
        return tuple(numpy.meshgrid(ky, kx, indexing='ij'))
