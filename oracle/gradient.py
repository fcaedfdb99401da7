"""Gradient engine restated in numpy (reference phaser/engines/gradient/run.py:156-438 and
gradient/solvers.py:48-138).  Test infrastructure only.

jax and optax are absent, so:
  * `model_loss` restates run_model's forward model and loss (run.py:357-405) exactly;
  * `model_grad` is an explicit adjoint of it in torch's convention (2 dL/dconj(z) for
    complex z, dL/dx for real x).  It is pinned by central finite differences of
    `model_loss` in FP64 (tests/test_oracle_gradient.py), not by jax (parity with the jax
    engine is unpinned);
  * the optax transforms (trace, scale_by_adam, scale_by_polyak, scale_by_learning_rate)
    are restated from optax's published update equations (parity with optax unpinned).
"""
import numpy

from .misc import GroupManager
from .num import fft2, ifft2, abs2, check_finite, to_complex_dtype
from .optics import fourier_shift_filter, make_propagators, tilt_propagators


def _cut(state, group, n):
    osamp = state.object.sampling
    group_scan = state.scan[tuple(group)]
    group_obj = osamp.get_view_at_pos(state.object.data, group_scan, (n, n))
    shifts = osamp.get_subpx_shifts(group_scan, (n, n))
    return group_scan, group_obj, shifts


def loss_from_intensity(noise, model_intensity, patterns, mask):
    """noise_models.py:29-45 (amplitude/Anscombe) and :86-102 (Poisson), FP of the inputs."""
    p = numpy.maximum(patterns, 0.0)
    if noise['name'] == 'poisson':
        eps = noise['eps']
        return numpy.sum(mask * (model_intensity + eps + p * (numpy.log(p + eps) - numpy.log(model_intensity + eps) - 1.0)))
    c, eps, var = noise['offset'], noise['eps'], 1 + noise['gaussian_variance']
    return 2. * numpy.sum(mask * (numpy.sqrt(p + c) - numpy.sqrt(model_intensity + c) - eps)**2) / var


def group_props(state, group, ky, kx, props):
    """The group's propagators, tilted per position when the state has tilts
    (run.py:373,391: tilt_propagators of the group's tilts): (S-1, B or 1, n, n)."""
    if props is None:
        return None
    tilts = state.tilt[tuple(group)] if state.tilt is not None else None
    return tilt_propagators(ky, kx, state.object.thicknesses, props, tilts)


def forward(state, group, ky, kx, props):
    """Shifted probes, per-slice waves, exit spectra (run.py:375-393)."""
    probe = state.probe.data
    n = probe.shape[-1]
    group_scan, group_obj, shifts = _cut(state, group, n)
    filt = fourier_shift_filter(ky, kx, shifts)[:, None, ...]
    psi = ifft2(fft2(probe) * filt)
    waves = [psi]
    S = group_obj.shape[1]
    tprops = group_props(state, group, ky, kx, props)
    for s in range(S - 1):
        psi = ifft2(fft2(psi * group_obj[:, s, None]) * tprops[s][:, None])
        waves.append(psi)
    model_wave = fft2(psi * group_obj[:, -1, None])
    return waves, group_obj, filt, model_wave


def model_loss(state, group, patterns, mask, noise, ky, kx, props):
    """Loss of run_model for one group (gradient/run.py:357-405), no regularizers."""
    _w, _o, _f, model_wave = forward(state, group, ky, kx, props)
    model_intensity = numpy.sum(abs2(model_wave), axis=1)
    return loss_from_intensity(noise, model_intensity, patterns[tuple(group)], mask)


def model_grad(state, group, patterns, mask, noise, ky, kx, props):
    """(loss, grad_object (S,H,W), grad_probe (M,n,n), grad_positions (B,2)), torch convention."""
    waves, group_obj, filt, Psi = forward(state, group, ky, kx, props)
    I = numpy.sum(abs2(Psi), axis=1)
    D = numpy.maximum(patterns[tuple(group)], 0.0)
    loss = loss_from_intensity(noise, I, patterns[tuple(group)], mask)
    if noise['name'] == 'poisson':
        dLdI = mask * (1.0 - D / (I + noise['eps']))
    else:
        c, eps, var = noise['offset'], noise['eps'], 1 + noise['gaussian_variance']
        si = numpy.sqrt(I + c)
        dLdI = -2.0 * mask * (numpy.sqrt(D + c) - si - eps) / (var * si)
    g = ifft2(2.0 * dLdI[:, None] * Psi)               # gradient w.r.t. the exit wave
    osamp = state.object.sampling
    group_scan = state.scan[tuple(group)]
    S = group_obj.shape[1]
    g_obj = numpy.zeros_like(state.object.data)
    n = state.probe.data.shape[-1]
    tprops = group_props(state, group, ky, kx, props)
    for s in range(S - 1, -1, -1):
        g_obj[s] = osamp.add_view_at_pos(g_obj[s], group_scan, numpy.sum(g * numpy.conj(waves[s]), axis=1))
        g_psi = g * numpy.conj(group_obj[:, s, None])
        if s > 0:
            g = ifft2(fft2(g_psi) * numpy.conj(tprops[s - 1])[:, None])
        else:
            g = g_psi
    Gk = fft2(g)                                         # spectrum of dL/dpsi0
    g_probe = numpy.sum(ifft2(Gk * numpy.conj(filt)), axis=0)
    Pk = fft2(state.probe.data)[None] * filt
    t = numpy.conj(Gk) * Pk
    g_pos = numpy.stack([numpy.sum(numpy.real(t * (-2j * numpy.pi * k)), axis=(1, -2, -1)) for k in (ky, kx)], axis=-1)
    return loss, g_obj, g_probe, g_pos


def dry_run(state, group, patterns, ky, kx, props):
    """gradient/run.py:408-438: sum(D) / sum |fft2(exit)|^2 (last slice, no mask)."""
    _w, _o, _f, model_wave = forward(state, group, ky, kx, props)
    return numpy.sum(patterns[tuple(group)], axis=(-2, -1)) / numpy.sum(abs2(model_wave), axis=(1, -2, -1))


class Solver:
    """optax transform chains of gradient/solvers.py:88-138 on numpy arrays."""

    def __init__(self, kind, params, **hp):
        self.kind, self.params, self.hp = kind, frozenset(params), hp
        self.state = {}
        self.count = 0

    def update(self, directions, loss):
        """directions: {name: d} steepest-descent directions; returns {name: update}."""
        self.count += 1
        out = {}
        if self.kind == 'polyak_sgd':
            u1 = {k: self.hp['scaling'] * d for k, d in directions.items()}
            sq = sum(float(numpy.sum(abs2(u))) for u in u1.values())
            eps = self.hp.get('eps', 0.0)
            step = 0.0 if sq + eps <= numpy.finfo(float).eps else \
                min((loss - self.hp['f_min']) / (sq + eps), self.hp['max_learning_rate'])
            return {k: step * u for k, u in u1.items()}
        for k, d in directions.items():
            if self.kind == 'sgd':
                m = self.hp.get('momentum')
                if m is not None:
                    t = d + m * self.state.get(k, numpy.zeros_like(d))
                    self.state[k] = t
                    u = d + m * t if self.hp.get('nesterov', True) else t
                else:
                    u = d
                out[k] = self.hp['learning_rate'] * u
            else:  # adam
                b1, b2 = self.hp.get('b1', 0.9), self.hp.get('b2', 0.999)
                mu, nu = self.state.get(k, (numpy.zeros_like(d), numpy.zeros(d.shape, dtype=abs2(d).dtype)))
                mu = (1 - b1) * d + b1 * mu
                nu = (1 - b2) * abs2(d) + b2 * nu
                self.state[k] = (mu, nu)
                c = self.count
                if self.hp.get('nesterov', False):
                    mh = b1 * mu / (1 - b1**(c + 1)) + (1 - b1) * d / (1 - b1**c)
                else:
                    mh = mu / (1 - b1**c)
                u = mh / (numpy.sqrt(nu / (1 - b2**c) + self.hp.get('eps_root', 0.0)) + self.hp.get('eps', 1e-8))
                out[k] = self.hp['learning_rate'] * u
        return out


def run_engine(state, patterns, mask, *, noise, solvers, niter, grouping=None, seed=None, bwlim_frac=2 / 3,
               dtype=numpy.float32, update_object=True, update_probe=True, update_positions=False,
               compact=False, shuffle_groups=None, regularizers=(), group_constraints=(), iter_constraints=()):
    """Gradient engine loop (run.py:156-290) with group solvers for object/probe and
    per-iteration solvers for positions.  `solvers`: list of oracle Solver.
    `regularizers`: oracle.regularizers cost regularizers, whose loss joins the group loss
    and whose object gradient joins the model's (run.py:399-403); constraints are applied
    after each group's solver steps (run.py:344-347) and after the per-iteration solvers
    (run.py:276-279)."""
    cdt = to_complex_dtype(dtype)
    ky, kx = state.probe.sampling.recip_grid(dtype=dtype)
    props = make_propagators(ky, kx, state.probe.sampling.k_max, state.wavelength, state.object.thicknesses, bwlim_frac)
    groups = GroupManager(state.scan, grouping, compact, seed)
    shuffle = shuffle_groups or not compact     # run.py:185
    factors = numpy.concatenate([dry_run(state, g, patterns, ky, kx, props) for g in groups.iter(state.scan)])
    state.probe.data *= numpy.sqrt(numpy.mean(factors))
    probe_int = numpy.sum(abs2(state.probe.data))
    group_solvers = [s for s in solvers if not (s.params <= {'positions', 'tilt'})]
    iter_solvers = [s for s in solvers if s.params <= {'positions', 'tilt'}]
    start_i = state.iter.total_iter
    gc_states = [c.init_state(state) for c in group_constraints]
    ic_states = [c.init_state(state) for c in iter_constraints]
    for i in range(1, niter + 1):
        state.iter.engine_iter = i
        losses = []
        pos_grads = numpy.zeros_like(state.scan)
        for gi, group in enumerate(groups.iter(state.scan, i, shuffle)):
            loss, g_obj, g_probe, g_pos = model_grad(state, group, patterns, mask, noise, ky, kx, props)
            for reg in regularizers:
                rl, _st = reg.calc_loss_group(group, state, None)
                loss = loss + rl
                g_obj = g_obj + reg.grad_group(group, state)
            B = group.shape[-1]
            dirs = {'object': (-g_obj / (probe_int * B)).astype(cdt), 'probe': (-g_probe / B).astype(cdt)}
            pos_grads[tuple(group)] = g_pos / (probe_int * B)
            for s in group_solvers:
                use = {k: v for k, v in dirs.items() if k in s.params and
                       ((k == 'object' and update_object) or (k == 'probe' and update_probe))}
                if use:
                    upd = s.update(use, float(loss))
                    if 'object' in upd:
                        state.object.data += upd['object'].astype(cdt)
                    if 'probe' in upd:
                        state.probe.data += upd['probe'].astype(cdt)
            for k, c in enumerate(group_constraints):
                _s, gc_states[k] = c.apply_group(group, state, gc_states[k])
            losses.append(loss)
            check_finite(state.object.data, state.probe.data, context=f"object or probe, group {gi}")
        loss = float(numpy.mean(numpy.array(losses, dtype=dtype)))
        if update_positions:
            for s in iter_solvers:
                upd = s.update({'positions': -pos_grads}, loss)['positions']
                upd = upd - numpy.mean(upd, tuple(range(upd.ndim - 1)))
                state.scan += upd.astype(state.scan.dtype)
        for k, c in enumerate(iter_constraints):
            _s, ic_states[k] = c.apply_iter(state, ic_states[k])
        state.progress.iters = numpy.concatenate([state.progress.iters, [i + start_i]])
        state.progress.detector_errors = numpy.concatenate([state.progress.detector_errors, [loss]])
    return state
