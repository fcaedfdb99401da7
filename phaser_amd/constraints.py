"""Constraints and cost regularizers on the MI355X backend (reference
phaser/engines/common/regularizers.py, protocols phaser/hooks/regularization.py:16-33).

Each constraint is built from its plan entry, prepares its device tables once
(`init_state`, outside any captured graph) and then launches its kernel on the group
workspace (`apply`): nothing is allocated and nothing synchronises, so group constraints
run inside the captured iteration.  The host tables are computed with the reference's
numpy expressions (masks, filters, taps) and uploaded, so the kernels see the same
constants.

  clamp_object_amplitude  phx_clamp_amplitude_c64    object, elementwise
  limit_probe_support     phx_probe_limit_support    probe modes, FFT in LDS
  layers                  phx_regularize_layers_c64  object, along the slice axis
  obj_low_pass            phx_object_filter_c64      object, Bluestein DFT per line
  obj_gaussian            phx_object_filter_c64
  obj_l1, obj_l2, obj_tikh, layers_tv, layers_tikh   phx_cost_regularizers_c64 (gradient
                                                     engine: loss and object gradient)
"""
import ctypes

import numpy

from . import _lib
from .device import stream_handle, to_device_scrambled
from .plan import (ClampObjectAmplitudePlan, CostRegularizerPlan, LimitProbeSupportPlan, ObjGaussianPlan,
                   ObjLowPassPlan, RegularizeLayersPlan, UnsupportedPlan)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class _Constraint:
    target = 'object'   # which state array the constraint rewrites

    def __init__(self, plan):
        self.plan = plan

    def init_state(self, ws, state):
        return self

    def apply(self, ws, g=None, flag_value=None, stream=None):
        raise NotImplementedError


class ClampObjectAmplitude(_Constraint):
    """regularizers.py:20-42."""

    def apply(self, ws, g=None, flag_value=None, stream=None):
        st = stream if stream is not None else stream_handle()
        flag = None if flag_value is None else _p(ws.finite_flag)
        _lib.call("phx_clamp_amplitude_c64", _p(ws.object), ws.object.numel(), float(numpy.float32(self.plan.amplitude)),
                  flag, int(flag_value or 0), st)


class LimitProbeSupport(_Constraint):
    """regularizers.py:45-64: the mask |k|^2 <= (max_angle 1e-3 / lambda)^2 on the probe's
    FP64 reciprocal grid (num.py:651-670), uploaded in scrambled order."""
    target = 'probe'

    def init_state(self, ws, state):
        import torch
        samp = state.probe.sampling
        ky = numpy.fft.fftfreq(int(samp.shape[0]), float(samp.sampling[0]))
        kx = numpy.fft.fftfreq(int(samp.shape[1]), float(samp.sampling[1]))
        ky, kx = numpy.meshgrid(ky, kx, indexing='ij')
        mask = (kx**2 + ky**2 <= (self.plan.max_angle * 1e-3 / state.wavelength)**2).astype(numpy.float32)
        self.mask_s = to_device_scrambled(torch, mask, ws.n, False, ws.device).reshape(-1)
        return self

    def apply(self, ws, g=None, flag_value=None, stream=None):
        st = stream if stream is not None else stream_handle()
        _lib.call("phx_probe_limit_support", ws.desc_refs[0 if g is None else g], _p(self.mask_s), st)


class RegularizeLayers(_Constraint):
    """regularizers.py:86-132 (iter constraint only, as in the reference)."""

    def init_state(self, ws, state):
        import torch
        th = numpy.asarray(state.object.thicknesses)
        self.active = len(th) >= 2
        if not self.active:
            return self
        spacing = numpy.mean(th)
        r = int(numpy.ceil(2. * self.plan.sigma / spacing))
        n = 2 * r + 1
        zs = ((numpy.arange(0, n) - (n - 1) // 2) * spacing).astype(numpy.float32)
        kernel = numpy.exp(-(zs / self.plan.sigma)**2 / 2.)
        kernel /= numpy.sum(kernel)
        self.radius = r
        self.taps = torch.from_numpy(kernel.astype(numpy.float32)).to(ws.device)
        return self

    def apply(self, ws, g=None, flag_value=None, stream=None):
        if not self.active:
            return
        st = stream if stream is not None else stream_handle()
        w = float(self.plan.weight)
        _lib.call("phx_regularize_layers_c64", _p(ws.object), ws.S, ws.H, ws.W, _p(self.taps), self.radius,
                  float(numpy.float32(w)), float(numpy.float32(1 - w)), st)


def bluestein_tables(n):
    """Tables of a length-n DFT by Bluestein's algorithm: chirp w[k] = exp(-i pi k^2 / n)
    (k^2 reduced mod 2n in integers, the phase in FP64), bhat = FFT_L(conj w, wrapped) / L
    and the FFT_L twiddles exp(-2 pi i j / L), j < L / 2, for L = 2^p >= 2n - 1."""
    L = 1
    while L < 2 * n - 1:
        L *= 2
    L = max(L, 2)
    k = numpy.arange(n, dtype=numpy.int64)
    chirp = numpy.exp(-1j * numpy.pi * ((k * k) % (2 * n)).astype(numpy.float64) / n)
    b = numpy.zeros(L, dtype=numpy.complex128)
    b[:n] = numpy.conj(chirp)
    if n > 1:
        b[L - n + 1:] = numpy.conj(chirp[1:])[::-1]
    bhat = numpy.fft.fft(b) / L
    tw = numpy.exp(-2j * numpy.pi * numpy.arange(L // 2) / L)
    return chirp.astype(numpy.complex64), bhat.astype(numpy.complex64), tw.astype(numpy.complex64), L


class _ObjectFilter(_Constraint):
    """O_s <- ifft2(F fft2(O_s)) for a real filter F (H, W) on the corner-origin grid."""

    def filter(self, state, shape):
        raise NotImplementedError

    def init_state(self, ws, state):
        import torch
        H, W = ws.H, ws.W
        filt = numpy.asarray(self.filter(state, (H, W)), dtype=numpy.float32)
        dev = ws.device
        self.filt_t = torch.from_numpy(numpy.ascontiguousarray(filt.T)).to(dev)
        self.tables = []
        for n in (H, W):
            c, b, t, L = bluestein_tables(n)
            self.tables.append((torch.from_numpy(c).to(dev), torch.from_numpy(b).to(dev), torch.from_numpy(t).to(dev), L))
        self.work = torch.empty_like(ws.object)
        return self

    def apply(self, ws, g=None, flag_value=None, stream=None):
        st = stream if stream is not None else stream_handle()
        (ch, bh, th, lh), (cw, bw, tw, lw) = self.tables
        _lib.call("phx_object_filter_c64", _p(ws.object), _p(self.work), ws.S, ws.H, ws.W, _p(self.filt_t),
                  _p(ch), _p(bh), _p(th), lh, _p(cw), _p(bw), _p(tw), lw, st)


class ObjLowPass(_ObjectFilter):
    """regularizers.py:135-161: keep |k|^2 <= max_freq^2 (k in 1/px, FP64 fftfreq)."""

    def filter(self, state, shape):
        ky = numpy.fft.fftfreq(shape[0], 1.0)
        kx = numpy.fft.fftfreq(shape[1], 1.0)
        ky, kx = numpy.meshgrid(ky, kx, indexing='ij')
        return ky**2 + kx**2 <= self.plan.max_freq**2


class ObjGaussian(_ObjectFilter):
    """regularizers.py:164-192: 1 - w (1 - exp(-2 pi^2 sigma^2 k^2)) in float32."""

    def filter(self, state, shape):
        sampling = numpy.asarray(state.object.sampling.sampling)
        ky = numpy.fft.fftfreq(shape[0], sampling[0]).astype(numpy.float32)
        kx = numpy.fft.fftfreq(shape[1], sampling[1]).astype(numpy.float32)
        ky, kx = numpy.meshgrid(ky, kx, indexing='ij')
        k2 = ky**2 + kx**2
        filt = numpy.exp(-(2 * numpy.pi**2 * self.plan.sigma**2) * k2).astype(numpy.float32)
        return (1. - self.plan.weight * (1. - filt)).astype(numpy.float32)


_CLASSES = {ClampObjectAmplitudePlan: ClampObjectAmplitude, LimitProbeSupportPlan: LimitProbeSupport,
            RegularizeLayersPlan: RegularizeLayers, ObjLowPassPlan: ObjLowPass, ObjGaussianPlan: ObjGaussian}


def make_constraints(plans, ws, state, what):
    """Device constraints for a plan's `group_constraints` or `iter_constraints` list."""
    out = []
    for p in plans:
        if isinstance(p, UnsupportedPlan) or type(p) not in _CLASSES:
            name = getattr(p, 'name', type(p).__name__)
            raise NotImplementedError(f"{what} '{name}' is not implemented by the MI355X backend")
        out.append(_CLASSES[type(p)](p).init_state(ws, state))
    return out


# ------------------------------------------------------------------ cost regularizers
_COST_SLOTS = {'obj_l1': 0, 'obj_l2': 1, 'obj_tikh': 2, 'layers_tv': 3, 'layers_tikh': 4}


class CostRegularizers:
    """The gradient engine's object cost regularizers (regularizers.py:195-360) as one
    pass per group: loss into the group's FP64 loss slot, gradient into grad_obj."""

    def __init__(self, plans, n_slices):
        self.costs = numpy.zeros(5, dtype=numpy.float64)
        for p in plans:
            if isinstance(p, UnsupportedPlan) or not isinstance(p, CostRegularizerPlan) or p.name not in _COST_SLOTS:
                name = getattr(p, 'name', type(p).__name__)
                raise NotImplementedError(f"regularizer '{name}' is not implemented by the MI355X backend")
            if p.name in ('layers_tv', 'layers_tikh') and n_slices < 2:
                continue   # zero loss for a single slice (regularizers.py:331-332, 353-354)
            self.costs[_COST_SLOTS[p.name]] += p.cost
        self.active = bool(numpy.any(self.costs != 0))

    def apply(self, ws, group_size, loss_ptr, want_grad, stream=None):
        """Add the loss of a group of `group_size` positions to *loss_ptr and (want_grad)
        its gradient to ws.grad_obj."""
        if not self.active:
            return
        st = stream if stream is not None else stream_handle()
        # cost scale group.shape[-1] / n_positions in float32 (regularizers.py:207)
        scale = numpy.float32(group_size / ws.P)
        coefs = (ctypes.c_float * 5)(*[float(numpy.float32(c) * scale) for c in self.costs])
        _lib.call("phx_cost_regularizers_c64", _p(ws.object), _p(ws.grad_obj) if want_grad else None, ws.S, ws.H,
                  ws.W, ctypes.cast(coefs, ctypes.c_void_p), loss_ptr, int(bool(want_grad)), st)
