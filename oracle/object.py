"""Object sampling, index math and cutouts (reference phaser/utils/object.py).
Test infrastructure only."""
import numpy

from .misc import create_rng
from .num import to_real_dtype


def random_phase_object(shape, sigma=1e-6, *, seed=None, dtype=None):
    """cos(phi) + i sin(phi), phi ~ N(0, sigma) (object.py:34-53)."""
    rng = create_rng(seed, 'random_phase_object')
    real_dtype = to_real_dtype(dtype) if dtype is not None else numpy.float64
    obj_angle = numpy.array(rng.normal(0., sigma, tuple(shape)), dtype=real_dtype)
    return numpy.cos(obj_angle) + numpy.sin(obj_angle) * 1.j


class ObjectSampling:
    """Object grid: shape, sampling, corner, scanned region (object.py:189-481)."""

    def __init__(self, shape, sampling, corner=None, region_min=None, region_max=None):
        self.shape = numpy.broadcast_to(numpy.asarray(shape).astype(numpy.int_), (2,))
        self.sampling = numpy.broadcast_to(numpy.asarray(sampling).astype(numpy.float64), (2,))
        self.region_min = None if region_min is None else numpy.broadcast_to(numpy.asarray(region_min).astype(numpy.float64), (2,))
        self.region_max = None if region_max is None else numpy.broadcast_to(numpy.asarray(region_max).astype(numpy.float64), (2,))
        if corner is None:
            corner = -self.extent / 2. + self.sampling / 2.
        else:
            corner = numpy.broadcast_to(numpy.asarray(corner).astype(numpy.float64), (2,))
        self.corner = corner

    @property
    def min(self):
        return self.corner

    @property
    def max(self):
        return (self.corner + (self.shape - 1) * self.sampling).astype(numpy.float64)

    @property
    def extent(self):
        return (self.shape * self.sampling).astype(numpy.float64)

    def __eq__(self, other):
        return (isinstance(other, ObjectSampling) and numpy.array_equal(self.shape, other.shape)
                and numpy.array_equal(self.sampling, other.sampling)
                and numpy.array_equal(self.corner, other.corner))

    @staticmethod
    def _scan_extent(scan_positions):
        scan_min = numpy.array(tuple(float(numpy.nanmin(scan_positions[..., i])) for i in range(2)))
        scan_max = numpy.array(tuple(float(numpy.nanmax(scan_positions[..., i])) for i in range(2)))
        return scan_min, scan_max

    @classmethod
    def from_scan(cls, scan_positions, sampling, pad=0):
        """n = ceil((2 pad + scan_max - scan_min)/dx) + 1, corner = scan_min - pad (object.py:246-255)."""
        sampling = numpy.asarray(sampling).astype(numpy.float64)
        pad = numpy.broadcast_to(pad, (2,)).astype(numpy.float64)
        scan_min, scan_max = cls._scan_extent(scan_positions)
        n = numpy.ceil((2. * pad + scan_max - scan_min) / sampling).astype(numpy.int_) + 1
        return cls((n[0], n[1]), sampling, scan_min - pad, scan_min, scan_max)

    def expand_to_scan(self, scan_positions, pad=0.):
        """Grow the grid so every position is at least `pad` from the edge (object.py:257-275)."""
        pad = numpy.broadcast_to(pad, (2,)).astype(numpy.float64)
        scan_min, scan_max = self._scan_extent(scan_positions)
        pad_min = numpy.ceil(numpy.maximum(0, self.min - scan_min + pad) / self.sampling).astype(numpy.int_)
        pad_max = numpy.ceil(numpy.maximum(0, scan_max - self.max + pad) / self.sampling).astype(numpy.int_)
        if numpy.all(pad_min == 0) and numpy.all(pad_max == 0):
            return self
        region_min = numpy.minimum(self.region_min, scan_min) if self.region_min is not None else None
        region_max = numpy.maximum(self.region_max, scan_max) if self.region_max is not None else None
        return self.__class__(tuple(self.shape + pad_min + pad_max), self.sampling,
                              self.corner - pad_min * self.sampling, region_min, region_max)

    def check_scan(self, scan_positions, pad=0.):
        """Raise if positions lie completely outside the object (object.py:277-289)."""
        pad = numpy.broadcast_to(pad, (2,)).astype(numpy.float64)
        obj_min = self.corner - pad
        obj_max = self.corner + self.extent + pad
        outside = ((scan_positions[..., 0] < obj_min[0]) | (scan_positions[..., 0] > obj_max[0]) |
                   (scan_positions[..., 1] < obj_min[1]) | (scan_positions[..., 1] > obj_max[1]))
        if (n_outside := int(numpy.sum(outside))):
            raise ValueError(f"{n_outside}/{outside.size} probe positions completely outside object")

    def _pos_to_object_idx(self, pos, cutout_shape):
        """(pos - corner)/dx - max(0, n-1)/2, in FP64 (object.py:291-301)."""
        pos = numpy.asarray(pos)
        shift = -numpy.maximum(0., (numpy.array(cutout_shape[-2:]) - 1.)) / 2.
        return ((pos - self.corner) / self.sampling + shift).astype(numpy.float64)

    def start_idxs(self, pos, cutout_shape):
        """numpy.round (half to even) of the FP64 index (object.py:493-495)."""
        return numpy.round(self._pos_to_object_idx(pos, cutout_shape)).astype(numpy.int_)

    def slice_at_pos(self, pos, cutout_shape):
        """Slices of the cutout nearest `pos` (object.py:303-322); asserts start >= 0."""
        idxs = self._pos_to_object_idx(pos, cutout_shape)
        start_i, start_j = map(int, numpy.round(idxs).astype(numpy.int64))
        assert start_i >= 0 and start_j >= 0
        return (slice(start_i, start_i + cutout_shape[-2]), slice(start_j, start_j + cutout_shape[-1]))

    def get_subpx_shifts(self, pos, cutout_shape):
        """(idx - round(idx)) * dx, in FP64 (object.py:324-331)."""
        pos = self._pos_to_object_idx(numpy.asarray(pos), cutout_shape)
        return (pos - numpy.round(pos)).astype(numpy.float64) * self.sampling

    def get_region_crop(self, pad=0.):
        """Slices covering the scanned region (object.py:367-378)."""
        pad = numpy.asarray(pad)
        if self.region_min is None:
            min_i, min_j = None, None
        else:
            min_i, min_j = numpy.ceil(self._pos_to_object_idx(self.region_min - pad, (0, 0))).astype(numpy.int_)
        if self.region_max is None:
            max_i, max_j = None, None
        else:
            max_i, max_j = numpy.floor(self._pos_to_object_idx(self.region_max + pad, (0, 0))).astype(numpy.int_) + 1
        return (slice(min_i, max_i), slice(min_j, max_j))

    def grid(self, *, dtype=numpy.float64):
        ys = numpy.linspace(self.min[0], self.max[0], self.shape[0], endpoint=True, dtype=dtype)
        xs = numpy.linspace(self.min[1], self.max[1], self.shape[1], endpoint=True, dtype=dtype)
        return tuple(numpy.meshgrid(ys, xs, indexing='ij'))

    def mpl_extent(self, center=True):
        shift = self.min - self.sampling / 2. * int(center)
        return (shift[1], self.extent[1] + shift[1], self.extent[0] + shift[0], shift[0])

    # cutouts: the numpy loop path of ObjectCutout (object.py:501-557)
    def get_view_at_pos(self, arr, pos, shape):
        pos = numpy.asarray(pos)
        out = numpy.empty((*pos.shape[:-1], *arr.shape[:-2], *shape[-2:]), dtype=arr.dtype)
        for idx in numpy.ndindex(pos.shape[:-1]):
            out[tuple(idx)] = arr[(Ellipsis, *self.slice_at_pos(pos[idx], shape))]
        return out

    def add_view_at_pos(self, arr, pos, view):
        """arr[cutout] += view, sequentially in position order (object.py:540-557).

        Note the reference copies `arr` first (`xp.array(arr)`, object.py:344-345)."""
        arr = numpy.array(arr)
        pos = numpy.asarray(pos)
        for idx in numpy.ndindex(pos.shape[:-1]):
            arr[(Ellipsis, *self.slice_at_pos(pos[idx], view.shape))] += view[idx]
        return arr

    def set_view_at_pos(self, arr, pos, view):
        arr = numpy.array(arr)
        pos = numpy.asarray(pos)
        for idx in numpy.ndindex(pos.shape[:-1]):
            arr[(Ellipsis, *self.slice_at_pos(pos[idx], view.shape))] = view[idx]
        return arr
