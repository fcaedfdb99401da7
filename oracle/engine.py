"""Conventional engine loop and position solvers (reference
phaser/engines/conventional/run.py:15-122, solvers.py:18-128 and 293-383,
phaser/engines/common/position_correction.py).  Test infrastructure only."""
import numpy

from .misc import GroupManager, mask_fraction_of_groups
from .num import check_finite, to_complex_dtype
from .optics import make_propagators
from .solvers import Sim, lsqml_dry_run, lsqml_run, epie_run


def flag_value(flag, engine_iter):
    """bool or dict(after, every, before) evaluated like types.py:125-146."""
    if isinstance(flag, bool):
        return flag
    after = flag.get('after', 0)
    every = flag.get('every', 1)
    before = flag.get('before', None)
    i = engine_iter
    return (before is None or i < before) and i > after and (i - after) % every == 0


class SteepestDescentPositionSolver:
    """position_correction.py:13-34 (note: max_step_size is set from step_size, :16)."""

    def __init__(self, step_size=1e-2, max_step_size=None):
        self.step_size = step_size
        self.max_step_size = step_size

    def init_state(self, sim):
        return None

    def perform_update(self, positions, gradients, state):
        update = self.step_size * gradients
        if self.max_step_size is not None:
            update_mag = numpy.linalg.norm(update, axis=-1, keepdims=True)
            update *= numpy.minimum(update_mag, self.max_step_size) / update_mag
        return update, state


class MomentumPositionSolver:
    """position_correction.py:37-61."""

    def __init__(self, step_size=1e-2, max_step_size=None, momentum=0.9):
        self.step_size = step_size
        self.max_step_size = max_step_size
        self.momentum = momentum

    def init_state(self, sim):
        return numpy.zeros_like(sim.scan)

    def perform_update(self, positions, gradients, state):
        update = self.step_size * gradients + self.momentum * state
        if self.max_step_size is not None:
            update_mag = numpy.linalg.norm(update, axis=-1, keepdims=True)
            update *= numpy.minimum(update_mag, self.max_step_size) / update_mag
        return update, update


def run_engine(state, patterns, pattern_mask, *, solver='lsqml', noise_model, niter=10, grouping=None,
               compact=False, shuffle_groups=None, update_probe=True, update_object=True,
               update_positions=False, calc_error=True, calc_error_fraction=0.1, bwlim_frac=2 / 3,
               beta_object=1.0, beta_probe=1.0, illum_reg_object=1e-2, illum_reg_probe=1e-2, gamma=1e-4,
               position_solver=None, seed=None, dtype=numpy.float32, group_callback=None, max_groups=None,
               group_constraints=(), iter_constraints=()):
    """Conventional engine (run.py:15-122) with the LSQ-ML (solvers.py:18-128) or ePIE
    (solvers.py:293-383) solver.  `state` is modified in place and returned.

    Solver hyperparameters are constants here (the reference evaluates schedules once
    per iteration, solvers.py:81-85).  `max_groups` truncates every group loop (used to
    time a bounded CPU sample); `group_callback(i, group)` is called after each group.
    `group_constraints` / `iter_constraints`: oracle.regularizers objects, applied after
    each group (solvers.py:109-113, after the finite check) and after the group loop
    (run.py:93, before the position update)."""
    assert patterns.dtype == dtype and pattern_mask.dtype == dtype
    assert state.object.data.dtype == to_complex_dtype(dtype)
    assert state.probe.data.dtype == to_complex_dtype(dtype)

    sim = Sim(state, noise_model, dtype)
    shuffle_flag = shuffle_groups or not compact   # run.py:32
    groups = GroupManager(state.scan, grouping, compact, seed=seed)
    calc_error_mask = mask_fraction_of_groups(len(groups), calc_error_fraction)
    pos_state = None if position_solver is None else position_solver.init_state(state)
    gc_states = [c.init_state(state) for c in group_constraints]   # simulation.py:127-132
    ic_states = [c.init_state(state) for c in iter_constraints]
    start_i = state.iter.total_iter

    (ky64, kx64) = state.probe.sampling.recip_grid(dtype=dtype)
    propagators = make_propagators(ky64, kx64, state.probe.sampling.k_max, state.wavelength,
                                   state.object.thicknesses, bwlim_frac)

    # presolve (solvers.py:36-63 / 306-329): rescale the probe intensity
    rescale = []
    for gi, group in enumerate(groups.iter(state.scan)):
        if max_groups is not None and gi >= max_groups:
            break
        rescale.append(lsqml_dry_run(sim, group, patterns[tuple(group)], pattern_mask=pattern_mask, props=propagators))
    rescale_factor = numpy.mean(numpy.concatenate(rescale, axis=0))
    state.probe.data *= numpy.sqrt(rescale_factor)

    for i in range(1, niter + 1):
        state.iter.engine_iter = i
        state.iter.total_iter = start_i + i
        iter_update_positions = flag_value(update_positions, i)
        iter_shuffle = flag_value(shuffle_flag, i)
        upd_obj = flag_value(update_object, i)
        upd_probe = flag_value(update_probe, i)
        calc_err = flag_value(calc_error, i)

        pos_update = numpy.zeros_like(state.scan, dtype=dtype)
        iter_errors = []
        for gi, group in enumerate(groups.iter(state.scan, i, iter_shuffle)):
            if max_groups is not None and gi >= max_groups:
                break
            group_calc_error = calc_err and calc_error_mask[gi]
            group_patterns = patterns[tuple(group)]
            if solver == 'lsqml':
                errors, gpu = lsqml_run(
                    sim, group, group_patterns, pattern_mask=pattern_mask, props=propagators,
                    beta_object=beta_object, beta_probe=beta_probe, update_object=upd_obj,
                    update_probe=upd_probe, update_position=iter_update_positions,
                    calc_error=group_calc_error, illum_reg_object=illum_reg_object,
                    illum_reg_probe=illum_reg_probe, gamma=gamma)
            else:
                # ePIE never forwards update_position / calc_error (solvers.py:358-366)
                errors, gpu = epie_run(
                    sim, group, group_patterns, pattern_mask=pattern_mask, props=propagators,
                    beta_object=beta_object, beta_probe=beta_probe, update_object=upd_obj,
                    update_probe=upd_probe)
            check_finite(state.object.data, state.probe.data, context=f"object or probe, group {gi}")
            for k, c in enumerate(group_constraints):   # simulation.py:137-146
                _s, gc_states[k] = c.apply_group(group, state, gc_states[k])
            if iter_update_positions:
                pos_update[tuple(group)] = gpu
            if group_calc_error:
                iter_errors.append(errors)
            if group_callback is not None:
                group_callback(gi, group)

        for k, c in enumerate(iter_constraints):   # run.py:93, simulation.py:148-157
            _s, ic_states[k] = c.apply_iter(state, ic_states[k])

        if iter_update_positions:
            if position_solver is None:
                raise ValueError("Updating positions with no PositionSolver specified")
            pos_update -= numpy.mean(pos_update, tuple(range(pos_update.ndim - 1)))
            pos_update, pos_state = position_solver.perform_update(state.scan, pos_update, pos_state)
            pos_update -= numpy.mean(pos_update, tuple(range(pos_update.ndim - 1)))
            state.scan += pos_update
            assert state.scan.dtype == dtype
            state.object.sampling.check_scan(state.scan, state.probe.sampling.extent / 2.)

        if len(iter_errors):
            error = float(numpy.nanmean(numpy.concatenate(iter_errors)))
            state.progress.iters = numpy.concatenate([state.progress.iters, [i + start_i]])
            state.progress.detector_errors = numpy.concatenate([state.progress.detector_errors, [error]])
    return state
