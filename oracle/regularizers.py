"""Constraints and cost regularizers restated in numpy (reference
phaser/engines/common/regularizers.py).  Test infrastructure only.

Classes keep the reference protocol: `init_state(sim)`, `apply_group(group, sim, state)`
and `apply_iter(sim, state)` for constraints (hooks/regularization.py:16-27), and
`calc_loss_group(group, sim, state)` for cost regularizers, plus `grad_group`, the
gradient of that loss with respect to the object in torch's convention
(2 dL/dconj(O)) for the gradient oracle, which has no autodiff.

Pins (tests/test_oracle_regularizers.py): closed forms (clamp keeps the phase and caps
the modulus; a probe inside the support and a band-limited object pass their filters
unchanged; weight 0 is the identity; a slice-constant object is a fixed point of the
layer filter), brute force (the layer convolution as an explicit loop over taps with edge
repetition, the object filters as explicit DFT sums on small grids), and central finite
differences for every `grad_group`.
"""
import numpy
import scipy.ndimage

from .num import abs2, fft2, ifft2, to_real_dtype


class ClampObjectAmplitude:
    """regularizers.py:20-42."""

    def __init__(self, amplitude=1.1):
        self.amplitude = amplitude

    def init_state(self, sim):
        return None

    def apply_group(self, group, sim, state):
        return self.apply_iter(sim, state)

    def apply_iter(self, sim, state):
        amp = to_real_dtype(sim.object.data.dtype)(self.amplitude)
        obj = sim.object.data
        with numpy.errstate(invalid='ignore', divide='ignore'):
            obj_amp = numpy.abs(obj)
            scale = numpy.minimum(obj_amp, amp) / obj_amp
        sim.object.data = obj * scale
        return sim, None


class LimitProbeSupport:
    """regularizers.py:45-64: the probe's spectrum is cut to |k| <= max_angle / lambda."""

    def __init__(self, max_angle):
        self.max_angle = max_angle

    def init_state(self, sim):
        (ky, kx) = sim.probe.sampling.recip_grid()
        return kx**2 + ky**2 <= (self.max_angle * 1e-3 / sim.wavelength)**2

    def apply_group(self, group, sim, state):
        return self.apply_iter(sim, state)

    def apply_iter(self, sim, state):
        sim.probe.data = ifft2(fft2(sim.probe.data) * state)
        return sim, state


class RegularizeLayers:
    """regularizers.py:86-132: Gaussian smoothing of log(O) along the slice axis."""

    def __init__(self, sigma=50.0, weight=0.9):
        self.sigma, self.weight = sigma, weight

    def init_state(self, sim):
        return None

    def taps(self, thicknesses, dtype):
        """The normalised Gaussian taps and their radius r (regularizers.py:103-111)."""
        layer_spacing = numpy.mean(thicknesses)
        r = int(numpy.ceil(2. * self.sigma / layer_spacing))
        n = 2 * r + 1
        zs = ((numpy.arange(0, n) - (n - 1) // 2) * layer_spacing).astype(dtype)
        kernel = numpy.exp(-(zs / self.sigma)**2 / 2.)
        kernel /= numpy.sum(kernel)
        return kernel, r

    def apply_iter(self, sim, state):
        if len(sim.object.thicknesses) < 2:
            return sim, None
        dtype = to_real_dtype(sim.object.data.dtype)
        kernel, _r = self.taps(sim.object.thicknesses, dtype)
        new_obj = numpy.exp(scipy.ndimage.convolve1d(numpy.log(sim.object.data), kernel, axis=0, mode='nearest'))
        assert new_obj.dtype == sim.object.data.dtype
        sim.object.data = self.weight * new_obj + (1 - self.weight) * sim.object.data
        return sim, None


class ObjLowPass:
    """regularizers.py:135-161: the object's spectrum cut to |k| <= max_freq (1/px)."""

    def __init__(self, max_freq=0.4):
        self.max_freq = max_freq

    def init_state(self, sim):
        shape = sim.object.data.shape[-2:]
        ky = numpy.fft.fftfreq(shape[0], 1.0)
        kx = numpy.fft.fftfreq(shape[1], 1.0)
        (ky, kx) = numpy.meshgrid(ky, kx, indexing='ij')
        return ky**2 + kx**2 <= self.max_freq**2

    def apply_group(self, group, sim, state):
        return self.apply_iter(sim, state)

    def apply_iter(self, sim, state):
        sim.object.data = ifft2(state * fft2(sim.object.data))
        return sim, state


class ObjGaussian:
    """regularizers.py:164-192: Gaussian blur of weight w in Fourier space."""

    def __init__(self, sigma, weight=0.9):
        self.sigma, self.weight = sigma, weight

    def init_state(self, sim):
        shape = sim.object.data.shape[-2:]
        sampling = sim.object.sampling.sampling
        dtype = to_real_dtype(sim.object.data.dtype)
        ky = numpy.fft.fftfreq(shape[0], sampling[0]).astype(dtype)
        kx = numpy.fft.fftfreq(shape[1], sampling[1]).astype(dtype)
        (ky, kx) = numpy.meshgrid(ky, kx, indexing='ij')
        k2 = ky**2 + kx**2
        filt = numpy.exp(-(2 * numpy.pi**2 * self.sigma**2) * k2).astype(dtype)
        return (1. - self.weight * (1. - filt)).astype(dtype)

    def apply_group(self, group, sim, state):
        return self.apply_iter(sim, state)

    def apply_iter(self, sim, state):
        sim.object.data = ifft2(state * fft2(sim.object.data))
        return sim, state


# ---------------------------------------------------------------- cost regularizers
def _cost_scale(group, sim, dtype):
    """group.shape[-1] / n_positions, as the reference computes it (regularizers.py:207)."""
    return (numpy.asarray(group).shape[-1] / numpy.prod(sim.scan.shape[:-1])).astype(dtype)


def _unit(d):
    a = numpy.abs(d)
    with numpy.errstate(invalid='ignore', divide='ignore'):
        return numpy.where(a > 0, d / numpy.where(a > 0, a, 1), 0)


class _CostRegularizer:
    def __init__(self, cost):
        self.cost = cost

    def init_state(self, sim):
        return None

    def calc_loss_group(self, group, sim, state):
        obj = sim.object.data
        cost = self._cost(obj)
        return cost * _cost_scale(group, sim, numpy.asarray(cost).dtype) * self.cost, state

    def grad_group(self, group, sim):
        """2 dL/dconj(O) of calc_loss_group's loss."""
        obj = sim.object.data
        scale = _cost_scale(group, sim, to_real_dtype(obj.dtype)) * self.cost
        return scale * self._grad(obj)


class ObjL1(_CostRegularizer):
    """regularizers.py:195-209; the gradient at O = 1 is taken as 0."""

    def _cost(self, obj):
        return numpy.sum(numpy.abs(obj - 1.0))

    def _grad(self, obj):
        return _unit(obj - 1.0)


class ObjL2(_CostRegularizer):
    """regularizers.py:212-226."""

    def _cost(self, obj):
        return numpy.sum(abs2(obj - 1.0))

    def _grad(self, obj):
        return 2. * (obj - 1.0)


class ObjTikhonov(_CostRegularizer):
    """regularizers.py:297-316."""

    def _cost(self, obj):
        return numpy.sum(abs2(numpy.diff(obj, axis=-1))) + numpy.sum(abs2(numpy.diff(obj, axis=-2)))

    def _grad(self, obj):
        g = numpy.zeros_like(obj)
        for ax in (-1, -2):
            d = numpy.diff(obj, axis=ax)
            pad_lo = [(0, 0)] * obj.ndim
            pad_hi = [(0, 0)] * obj.ndim
            pad_lo[ax] = (1, 0)
            pad_hi[ax] = (0, 1)
            # dL/dconj(O_j) = d_{j-1} - d_j with d = O_{j+1} - O_j
            g += 2. * (numpy.pad(d, pad_lo) - numpy.pad(d, pad_hi))
        return g


class LayersTotalVariation(_CostRegularizer):
    """regularizers.py:319-338; the gradient of |d| at d = 0 is taken as 0."""

    def calc_loss_group(self, group, sim, state):
        if sim.object.data.shape[0] < 2:
            return 0.0, state
        return super().calc_loss_group(group, sim, state)

    def _cost(self, obj):
        return numpy.sum(numpy.abs(numpy.diff(obj, axis=0)))

    def _grad(self, obj):
        if obj.shape[0] < 2:
            return numpy.zeros_like(obj)
        u = _unit(numpy.diff(obj, axis=0))
        return numpy.pad(u, ((1, 0), (0, 0), (0, 0))) - numpy.pad(u, ((0, 1), (0, 0), (0, 0)))


class LayersTikhonov(_CostRegularizer):
    """regularizers.py:341-360."""

    def calc_loss_group(self, group, sim, state):
        if sim.object.data.shape[0] < 2:
            return 0.0, state
        return super().calc_loss_group(group, sim, state)

    def _cost(self, obj):
        return numpy.sum(abs2(numpy.diff(obj, axis=0)))

    def _grad(self, obj):
        if obj.shape[0] < 2:
            return numpy.zeros_like(obj)
        d = numpy.diff(obj, axis=0)
        return 2. * (numpy.pad(d, ((1, 0), (0, 0), (0, 0))) - numpy.pad(d, ((0, 1), (0, 0), (0, 0))))


CONSTRAINTS = {
    'clamp_object_amplitude': ClampObjectAmplitude, 'limit_probe_support': LimitProbeSupport,
    'layers': RegularizeLayers, 'obj_low_pass': ObjLowPass, 'obj_gaussian': ObjGaussian,
}
REGULARIZERS = {
    'obj_l1': ObjL1, 'obj_l2': ObjL2, 'obj_tikh': ObjTikhonov, 'obj_tikhonov': ObjTikhonov,
    'layers_tv': LayersTotalVariation, 'layers_tikh': LayersTikhonov, 'layers_tikhonov': LayersTikhonov,
}


def make(spec, known):
    """A constraint or regularizer from its plan form {'type': name, **props}."""
    d = dict(spec)
    return known[d.pop('type')](**d)
