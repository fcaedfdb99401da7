"""Host-side sampling grids and cutout index math.

Start indices and subpixel shifts are computed here in FP64 numpy with the exact
expression order of phaser/utils/object.py:291-331 (half-to-even rounding of
(pos - corner)/dx - max(0, n-1)/2), then uploaded; the kernels never recompute them, so
jitter-free raster scans that land exactly on .5 round the same way as the reference.
"""
import numpy


class Sampling:
    """Probe / pattern grid (phaser/utils/num.py:544-670)."""

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

    def __eq__(self, other):
        return (isinstance(other, Sampling) and numpy.array_equal(self.shape, other.shape)
                and numpy.array_equal(self.extent, other.extent))

    @property
    def k_max(self):
        return (1 / (2 * self.sampling)).astype(numpy.float64)

    def recip_1d(self, dtype=numpy.float32):
        """fftfreq along y and x cast to `dtype` (corner origin)."""
        ky = numpy.fft.fftfreq(int(self.shape[0]), float(self.sampling[0])).astype(dtype)
        kx = numpy.fft.fftfreq(int(self.shape[1]), float(self.sampling[1])).astype(dtype)
        return ky, kx

    def recip_grid(self, *, dtype=numpy.float32):
        ky, kx = self.recip_1d(dtype)
        return tuple(numpy.meshgrid(ky, kx, indexing='ij'))

    def resample(self, arr, new_samp, *, order=1, cval=0.0):
        """Resample real-space `arr` (..., ny, nx) onto `new_samp`, centres aligned
        (num.py:712-734; linear interpolation, zeros outside)."""
        matrix = new_samp.sampling / self.sampling
        offset = ((self.shape - 1) - matrix * (new_samp.shape - 1)) / 2.
        return affine_transform(arr, matrix, offset, tuple(int(v) for v in new_samp.shape), order=order, cval=cval)

    def resample_recip(self, arr, new_samp, *, order=1, cval=0.0):
        """Resample corner-origin reciprocal-space `arr` onto `new_samp`'s reciprocal grid
        (num.py:736-754): fftshift, resample as a grid of extent k_max, ifftshift."""
        old = Sampling(self.shape, extent=tuple(self.k_max))
        new = Sampling(new_samp.shape, extent=tuple(new_samp.k_max))
        arr = numpy.fft.fftshift(arr, axes=(-1, -2))
        return numpy.fft.ifftshift(old.resample(arr, new, order=order, cval=cval), axes=(-1, -2))


def affine_transform(arr, matrix, offset, output_shape, *, order=1, cval=0.0):
    """scipy.ndimage.affine_transform over the last two axes of `arr`, plane by plane,
    mode 'grid-constant' (utils/image.py:151-195); complex planes part by part."""
    import warnings
    import scipy.ndimage
    arr = numpy.asarray(arr)
    lead = arr.shape[:-2]
    flat = arr.reshape(-1, *arr.shape[-2:])
    out = numpy.empty((flat.shape[0], *output_shape), dtype=arr.dtype)
    with warnings.catch_warnings():
        # a 1-D `matrix` is the diagonal (a scale per axis), as intended
        warnings.filterwarnings('ignore', message="The behavior of affine_transform with a 1-D array")
        for i in range(flat.shape[0]):
            out[i] = scipy.ndimage.affine_transform(flat[i], numpy.asarray(matrix), offset=offset,
                                                    output_shape=output_shape, order=order, mode='grid-constant',
                                                    cval=cval)
    return out.reshape(*lead, *output_shape)


class ObjectSampling:
    """Object grid (phaser/utils/object.py:189-481), the parts the engines use."""

    def __init__(self, shape, sampling, corner=None, region_min=None, region_max=None):
        self.shape = numpy.broadcast_to(numpy.asarray(shape).astype(numpy.int_), (2,))
        self.sampling = numpy.broadcast_to(numpy.asarray(sampling).astype(numpy.float64), (2,))
        self.region_min = None if region_min is None else numpy.broadcast_to(numpy.asarray(region_min).astype(numpy.float64), (2,))
        self.region_max = None if region_max is None else numpy.broadcast_to(numpy.asarray(region_max).astype(numpy.float64), (2,))
        if corner is None:
            corner = -self.extent / 2. + self.sampling / 2.
        self.corner = numpy.broadcast_to(numpy.asarray(corner).astype(numpy.float64), (2,))

    @property
    def extent(self):
        return (self.shape * self.sampling).astype(numpy.float64)

    @property
    def min(self):
        return self.corner

    @property
    def max(self):
        return (self.corner + (self.shape - 1) * self.sampling).astype(numpy.float64)

    def __eq__(self, other):
        return (isinstance(other, ObjectSampling) and numpy.array_equal(self.shape, other.shape)
                and numpy.array_equal(self.sampling, other.sampling) and numpy.array_equal(self.corner, other.corner))

    @staticmethod
    def _scan_extent(scan):
        scan_min = numpy.array([float(numpy.nanmin(scan[..., i])) for i in range(2)])
        scan_max = numpy.array([float(numpy.nanmax(scan[..., i])) for i in range(2)])
        return scan_min, scan_max

    @classmethod
    def from_scan(cls, scan, sampling, pad=0):
        """object.py:246-255"""
        sampling = numpy.asarray(sampling).astype(numpy.float64)
        pad = numpy.broadcast_to(pad, (2,)).astype(numpy.float64)
        scan_min, scan_max = cls._scan_extent(scan)
        n = numpy.ceil((2. * pad + scan_max - scan_min) / sampling).astype(numpy.int_) + 1
        return cls((n[0], n[1]), sampling, scan_min - pad, scan_min, scan_max)

    def expand_to_scan(self, scan, pad=0.):
        """Grow the grid (corner moves, sampling kept) until it covers the scan plus `pad`
        (object.py:257-275)."""
        pad = numpy.broadcast_to(pad, (2,)).astype(numpy.float64)
        scan_min, scan_max = self._scan_extent(numpy.asarray(scan))
        pad_min = numpy.ceil(numpy.maximum(0, self.min - scan_min + pad) / self.sampling).astype(numpy.int_)
        pad_max = numpy.ceil(numpy.maximum(0, scan_max - self.max + pad) / self.sampling).astype(numpy.int_)
        if numpy.all(pad_min == 0) and numpy.all(pad_max == 0):
            return self
        region_min = numpy.minimum(self.region_min, scan_min) if self.region_min is not None else None
        region_max = numpy.maximum(self.region_max, scan_max) if self.region_max is not None else None
        return ObjectSampling(tuple(self.shape + pad_min + pad_max), self.sampling,
                              self.corner - pad_min * self.sampling, region_min, region_max)

    def with_sampling(self, sampling):
        """The same region at another pixel size, grown to whole pixels (object.py:469-481)."""
        sampling = numpy.broadcast_to(numpy.asarray(sampling, dtype=numpy.float64), (2,))
        new_shape = numpy.ceil(self.sampling / sampling * self.shape).astype(int)
        growth = sampling * new_shape - self.sampling * self.shape
        return ObjectSampling(tuple(new_shape), sampling, self.corner - growth / 2. + (sampling - self.sampling) / 2.,
                              self.region_min, self.region_max)

    def resample(self, arr, new_samp, *, order=1, cval=1.0):
        """Object values on `new_samp`'s grid (object.py:430-467 without rotation): linear
        interpolation, `cval` (vacuum, 1) outside."""
        matrix = new_samp.sampling / self.sampling
        offset = (new_samp.corner - self.corner) / self.sampling
        return affine_transform(arr, matrix, offset, tuple(int(v) for v in new_samp.shape), order=order, cval=cval)

    def check_scan(self, scan, pad=0.):
        """Raise if positions lie completely outside the object (object.py:277-289)."""
        pad = numpy.broadcast_to(pad, (2,)).astype(numpy.float64)
        lo = self.corner - pad
        hi = self.corner + self.extent + pad
        out = ((scan[..., 0] < lo[0]) | (scan[..., 0] > hi[0]) | (scan[..., 1] < lo[1]) | (scan[..., 1] > hi[1]))
        if (n := int(numpy.sum(out))):
            raise ValueError(f"{n}/{out.size} probe positions completely outside object")

    def pos_to_object_idx(self, pos, cutout_shape):
        """(pos - corner)/dx - max(0, n-1)/2 in FP64 (object.py:291-301)."""
        shift = -numpy.maximum(0., (numpy.array(cutout_shape[-2:]) - 1.)) / 2.
        return ((numpy.asarray(pos) - self.corner) / self.sampling + shift).astype(numpy.float64)

    def cutout_indices(self, pos, cutout_shape):
        """(start (..., 2) int64, subpixel shift (..., 2) float64 in length units).

        start = numpy.round (half to even) of the index (object.py:493-495); shift =
        (idx - round(idx)) * dx (object.py:324-331)."""
        idx = self.pos_to_object_idx(pos, cutout_shape)
        r = numpy.round(idx)
        return r.astype(numpy.int64), ((idx - r).astype(numpy.float64) * self.sampling)

    def get_region_crop(self, pad=0.):
        """object.py:367-378"""
        pad = numpy.asarray(pad)
        if self.region_min is None:
            min_i = min_j = None
        else:
            min_i, min_j = numpy.ceil(self.pos_to_object_idx(self.region_min - pad, (0, 0))).astype(numpy.int_)
        if self.region_max is None:
            max_i = max_j = None
        else:
            max_i, max_j = numpy.floor(self.pos_to_object_idx(self.region_max + pad, (0, 0))).astype(numpy.int_) + 1
        return (slice(min_i, max_i), slice(min_j, max_j))
