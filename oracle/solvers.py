"""LSQ-ML and ePIE group steps and dry runs (reference
phaser/engines/conventional/solvers.py, phaser/engines/common/simulation.py:241-306).
Test infrastructure only.

`sim` is an `oracle.solvers.Sim`: the reconstruction state plus the probe k-grid in the
simulation dtype and the noise model, as in SimulationState (simulation.py:86-135).
"""
import numpy

from .num import fft2, ifft2, abs2
from .optics import fourier_shift_filter, tilt_propagators


class Sim:
    """SimulationState subset (simulation.py:86-135)."""

    def __init__(self, state, noise_model, dtype):
        self.state = state
        self.noise_model = noise_model
        self.dtype = dtype
        (self.ky, self.kx) = state.probe.sampling.recip_grid(dtype=dtype)


def cutout_group(ky, kx, state, group, return_filters=False):
    """Object cutouts, subpixel filters and shifted probes (simulation.py:241-260)."""
    probes = state.probe.data
    group_scan = state.scan[tuple(group)]
    group_obj = state.object.sampling.get_view_at_pos(state.object.data, group_scan, probes.shape[-2:])
    filt = fourier_shift_filter(ky, kx, state.object.sampling.get_subpx_shifts(group_scan, probes.shape[-2:]))[:, None, ...]
    shifted = ifft2(fft2(probes) * filt)
    if return_filters:
        return shifted, group_obj, group_scan, filt
    return shifted, group_obj, group_scan


def _props_for(sim, props, group):
    st = sim.state
    return tilt_propagators(sim.ky, sim.kx, st.object.thicknesses, props,
                            st.tilt[tuple(group)] if st.tilt is not None else None)


def slice_forwards(props, state, f):
    """simulation.py:263-285 (numpy branch)."""
    if props is None:
        return f(0, None, state)
    n_slices = len(props) + 1
    for slice_i in range(n_slices - 1):
        state = f(slice_i, props[slice_i], state)
    return f(n_slices - 1, None, state)


def slice_backwards(props, state, f):
    """simulation.py:288-306 (numpy branch)."""
    if props is None:
        return f(0, None, state)
    n_slices = len(props) + 1
    for slice_i in range(n_slices - 1, 0, -1):
        state = f(slice_i, props[slice_i - 1], state)
    return f(0, None, state)


def lsqml_dry_run(sim, group, group_patterns, *, pattern_mask, props):
    """Rescale factors exp/model per position (solvers.py:131-167).

    Quirk kept: the wave is never multiplied by the last slice (solvers.py:147-156).
    The obj_mag / probe_mag accumulations of the reference are never read and are
    omitted."""
    psi, group_obj, _group_scan = cutout_group(sim.ky, sim.kx, sim.state, group)

    def run_slice(slice_i, prop, psi):
        if prop is not None:
            psi = ifft2(fft2(psi * group_obj[:, slice_i, None]) * prop[:, None])
        return psi

    psi = slice_forwards(_props_for(sim, props, group), psi, run_slice)
    model_intensity = numpy.sum(abs2(fft2(psi)), axis=(1, -2, -1))
    exp_intensity = numpy.sum(group_patterns * pattern_mask, axis=(-2, -1))
    return exp_intensity / model_intensity


epie_dry_run = lsqml_dry_run  # solvers.py:386-412 computes the same rescale factors


def _pos_step(sim, probes, group_obj, chi):
    """Position gradient step per position, (B, 2) (solvers.py:275-288)."""
    probes_fft = fft2(probes)
    probes_fft /= numpy.sum(abs2(probes), axis=(1, -1, -2), keepdims=True)

    def calc_pos_step(k):
        delta_P_x = ifft2(probes_fft * -2.j * numpy.pi * k)
        prod = delta_P_x * group_obj[:, 0, None]
        return numpy.sum(numpy.real(chi * numpy.conj(prod)), axis=(1, -1, -2)) / numpy.sum(abs2(prod))

    return numpy.stack(tuple(calc_pos_step(k) for k in (sim.ky, sim.kx)), axis=-1)


def lsqml_run(sim, group, group_patterns, *, pattern_mask, props,
              beta_object=0.9, beta_probe=0.9, update_object=True, update_probe=True,
              update_position=True, calc_error=True, illum_reg_object, illum_reg_probe, gamma):
    """One LSQ-ML group step, in place on sim.state (solvers.py:171-290).

    Returns (errors or None, pos_update or None).  The new_obj_mag/new_probe_mag
    accumulators of the reference are never read by anything and are omitted."""
    st = sim.state
    obj_grid = st.object.sampling
    n_slices = st.object.data.shape[0]
    eps = 1e-16

    probes, group_obj, group_scan, subpx_filters = cutout_group(sim.ky, sim.kx, st, group, return_filters=True)
    psi = numpy.zeros((n_slices, *probes.shape), dtype=probes.dtype)
    psi[0] = probes

    group_probe_mag = numpy.zeros(st.object.data.shape, dtype=sim.dtype)
    group_obj_mag = numpy.sum(abs2(numpy.prod(group_obj, axis=1)), axis=0)

    def sim_slice(slice_i, prop, state):
        gpm, psi = state
        gpm[slice_i] = obj_grid.add_view_at_pos(gpm[slice_i], group_scan, numpy.sum(abs2(psi[slice_i]), axis=1))
        if prop is not None:
            psi[slice_i + 1] = ifft2(fft2(psi[slice_i] * group_obj[:, slice_i, None]) * prop[:, None])
        return gpm, psi

    props = _props_for(sim, props, group)
    group_probe_mag, psi = slice_forwards(props, (group_probe_mag, psi), sim_slice)

    model_wave = fft2(psi[-1] * group_obj[:, -1, None])
    model_intensity = numpy.sum(abs2(model_wave), axis=1, keepdims=True)
    errors = numpy.sqrt(numpy.nansum((model_intensity - group_patterns[:, None])**2, axis=(1, -1, -2))) if calc_error else None

    chi = sim.noise_model.calc_wave_update(model_wave, model_intensity, group_patterns[:, None], pattern_mask)
    chi = ifft2(chi)

    def update_slice(slice_i, prop, chi):
        delta_P = chi * numpy.conj(group_obj[:, slice_i, None])
        if update_object:
            delta_O = chi * numpy.conj(psi[slice_i])
            alpha_O = numpy.sum(numpy.sum(numpy.real(chi * numpy.conj(delta_O * psi[slice_i])), axis=(-1, -2), keepdims=True), axis=1) \
                / (numpy.sum(abs2(delta_O * psi[slice_i])) + gamma)
            delta_O_avg = numpy.zeros_like(st.object.data[0])
            delta_O_avg = obj_grid.add_view_at_pos(delta_O_avg, group_scan, numpy.sum(delta_O, axis=1))
            delta_O_avg /= (group_probe_mag[slice_i] + illum_reg_object)
            obj_update = beta_object * numpy.sum(alpha_O * delta_O_avg * group_probe_mag[slice_i], axis=0) / (group_probe_mag[slice_i] + eps)
            st.object.data[slice_i] += obj_update

        if prop is not None:
            chi = ifft2(fft2(delta_P) * prop.conj()[:, None])
        elif update_probe:
            delta_P_avg = ifft2(numpy.sum(fft2(delta_P) * subpx_filters.conj(), axis=0))
            delta_P_avg /= (group_obj_mag + illum_reg_probe)
            alpha_P = numpy.sum(numpy.real(chi * numpy.conj(delta_P * group_obj[:, slice_i, None])), axis=(-1, -2), keepdims=True) \
                / (numpy.sum(abs2(delta_P * group_obj[:, slice_i, None])) + gamma)
            probe_update = beta_probe * numpy.sum(alpha_P * delta_P_avg * group_obj_mag, axis=0) / (group_obj_mag + eps)
            st.probe.data += probe_update
        return chi

    chi = slice_backwards(props, chi, update_slice)

    pos_update = _pos_step(sim, probes, group_obj, chi) if update_position else None
    return errors, pos_update


def epie_run(sim, group, group_patterns, *, pattern_mask, props, beta_object=0.9, beta_probe=0.9,
             update_object=True, update_probe=True, update_position=True, calc_error=True):
    """One ePIE group step, in place on sim.state (solvers.py:415-504)."""
    st = sim.state
    obj_grid = st.object.sampling
    n_slices = st.object.data.shape[0]

    probes, group_obj, group_scan, subpx_filters = cutout_group(sim.ky, sim.kx, st, group, return_filters=True)
    psi = numpy.zeros((n_slices, *probes.shape), dtype=probes.dtype)
    psi[0] = probes

    def sim_slice(slice_i, prop, psi):
        if prop is not None:
            psi[slice_i + 1] = ifft2(fft2(psi[slice_i] * group_obj[:, slice_i, None]) * prop[:, None])
        return psi

    props = _props_for(sim, props, group)
    psi = slice_forwards(props, psi, sim_slice)

    model_wave = fft2(psi[-1] * group_obj[:, -1, None])
    model_intensity = numpy.sum(abs2(model_wave), axis=1, keepdims=True)
    errors = numpy.sqrt(numpy.nansum((model_intensity - group_patterns[:, None])**2, axis=(1, -1, -2))) if calc_error else None
    chi = sim.noise_model.calc_wave_update(model_wave, model_intensity, group_patterns[:, None], pattern_mask)
    chi = ifft2(chi)

    def update_slice(slice_i, prop, chi):
        probe_update = group_obj[:, slice_i, None].conj() * chi / numpy.max(
            abs2(group_obj[:, slice_i, None]), axis=(-1, -2), keepdims=True)
        if update_object:
            group_obj_update = beta_object / n_slices * numpy.sum(psi[slice_i].conj() * chi, axis=1) / numpy.max(
                numpy.sum(abs2(psi[slice_i]), axis=1), axis=(-1, -2), keepdims=True)
            st.object.data[slice_i] = obj_grid.add_view_at_pos(st.object.data[slice_i], group_scan, group_obj_update)
        if prop is not None:
            chi = ifft2(fft2(probe_update) * prop.conj()[:, None])
        elif update_probe:
            probe_update = ifft2(numpy.mean(fft2(probe_update) * subpx_filters.conj(), axis=0))
            st.probe.data += beta_probe * probe_update
        return chi

    chi = slice_backwards(props, chi, update_slice)
    pos_update = _pos_step(sim, probes, group_obj, chi) if update_position else None
    return errors, pos_update
