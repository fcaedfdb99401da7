"""Conventional engines (LSQ-ML, ePIE) on the MI355X backend.

`run_engine(args, props)` has the signature, argument meaning, observer call sequence and
ProgressState updates of phaser/engines/conventional/run.py:15-122, and the solvers keep
the ConventionalSolver interface (phaser/hooks/solver.py:101-140): `init`, `presolve`,
`run_iteration(sim, groups, *, patterns, pattern_mask, propagators, update_object,
update_probe, update_positions, calc_error, calc_error_mask, observer)`.

Differences of mechanism, not of result:
  * every group step runs as HIP kernels on HBM-resident state (device.GroupWorkspace);
  * NaN/inf is detected inside the object/probe update kernels and reported once per
    iteration (ValueError naming the first bad group) instead of by a host sync after
    every group (solvers.py:109);
  * an iteration's launches can be captured once in a HIP graph and replayed.
"""
import logging

import numpy

from ..constraints import make_constraints
from ..device import GroupWorkspace, host_patterns
from ..groups import GroupManager, mask_fraction_of_groups
from ..optics import make_propagators
from ..plan import process_flag, process_schedule, shuffle_groups_flag
from .position import make_position_solver

logger = logging.getLogger(__name__)


class Observer:
    """No-op observer with the reference's callback names (phaser/observer.py:18-69)."""

    def init_engine(self, state, *, recons_name, plan, **kwargs): pass
    def start_engine(self, state): pass
    def update_group(self, state, force=False): pass
    def update_iteration(self, state, i, n, error=None): pass
    def finish_engine(self, state): pass


def wants_group_updates(observer):
    """Whether `observer` needs `update_group` after every group with that group's state
    (its own `wants_groups`, else whether it overrides this module's no-op)."""
    if hasattr(observer, 'wants_groups'):
        return bool(observer.wants_groups)
    fn = getattr(type(observer), 'update_group', None)
    return fn is not None and fn is not Observer.update_group


class SimulationState:
    """State plus device workspace (simulation.py:86-157)."""

    def __init__(self, *, state, noise_model, workspace, dtype=numpy.float32):
        self.state = state
        self.noise_model = noise_model
        self.workspace = workspace
        self.dtype = dtype


class _GraphCache:
    """One captured HIP graph per (flags) combination of an iteration's launches."""

    def __init__(self, enabled):
        self.enabled = enabled
        self.graphs = {}

    def run(self, key, launch):
        if not self.enabled:
            launch()
            return
        import torch
        g = self.graphs.get(key)
        if g is None:
            # warm-up outside capture, as torch requires for first-use side effects
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                launch()
            self.graphs[key] = g
        g.replay()


class _SolverBase:
    def __init__(self, engine_plan, props, use_graphs=True):
        self.engine_plan = engine_plan
        self.plan = props
        self.graphs = _GraphCache(use_graphs)
        # device group constraints (phaser_amd.constraints), applied after every group
        self.group_constraints = ()

    def init(self, sim):
        return sim

    def _schedule(self, name, sim):
        return process_schedule(getattr(self.plan, name))({'state': sim.state, 'niter': self.engine_plan.niter})

    def presolve(self, sim, groups, *, patterns, pattern_mask, propagators):
        """Rescale the probe intensity by sqrt(mean(exp/model)) (solvers.py:36-63, 306-329)."""
        ws = sim.workspace
        groups = list(groups)
        ws.set_groups(groups)
        for g in range(len(groups)):
            ws.launch_dry_run(g)
        errs = ws.errors.cpu().numpy()
        factors = numpy.concatenate([errs[g, :ws.sizes_host[g]] for g in range(len(groups))]).astype(numpy.float32)
        rescale = numpy.mean(factors)
        logger.info(f"Rescaling initial probe intensity by {rescale:.2e}")
        ws.scale_probe(numpy.sqrt(rescale))
        return sim

    def _launch_group(self, ws, g, **flags):
        raise NotImplementedError

    def run_iteration(self, sim, groups, *, patterns, pattern_mask, propagators, update_object, update_probe,
                      update_positions, calc_error, calc_error_mask, observer):
        ws = sim.workspace
        groups = list(groups)
        ws.set_groups(groups)
        self._set_params(sim)
        ws.reset_finite_flag()
        G = len(groups)
        err_groups = [g for g in range(G) if calc_error and calc_error_mask[g]]

        def launch_group(g):
            self._launch_group(ws, g, update_object=update_object, update_probe=update_probe,
                               update_position=update_positions, calc_error=(g in err_groups))

        if wants_group_updates(observer):
            # the observer sees the state after every group (solvers.py:109-119): groups run
            # one by one, each followed by the finite check and the callback
            ws.check_state_finite(0)
            for g in range(G):
                launch_group(g)
                bad = ws.finite_group()
                if bad is not None:
                    raise ValueError(f"NaN or inf encountered in object or probe, group {bad}")
                observer.update_group(sim.state, self.engine_plan.send_every_group)
        else:
            def launch():
                ws.check_state_finite(0)
                for g in range(G):
                    launch_group(g)

            key = (bool(update_object), bool(update_probe), bool(update_positions), bool(calc_error))
            self.graphs.run(key, launch)
            # the update kernels record the first group whose object or probe turned
            # non-finite; the check is made once the iteration's launches have run
            bad = ws.finite_group()
            if bad is not None:
                raise ValueError(f"NaN or inf encountered in object or probe, group {bad}")

        if update_positions:
            pos_update = ws.pos_update.cpu().numpy().reshape(*ws.scan_shape, 2)
        else:
            pos_update = numpy.zeros((*ws.scan_shape, 2), dtype=numpy.float32)
        iter_errors = []
        if err_groups:
            errs = ws.errors.cpu().numpy()
            iter_errors = [errs[g, :ws.sizes_host[g]] for g in err_groups]
        return sim, pos_update, iter_errors


class LSQMLSolver(_SolverBase):
    """phaser/engines/conventional/solvers.py:18-128"""

    @classmethod
    def name(cls):
        return "LSQML"

    def _set_params(self, sim):
        sim.workspace.set_params(
            beta_object=self._schedule('beta_object', sim), beta_probe=self._schedule('beta_probe', sim),
            reg_object=self._schedule('illum_reg_object', sim), reg_probe=self._schedule('illum_reg_probe', sim),
            gamma=self._schedule('gamma', sim))

    def _launch_group(self, ws, g, **flags):
        ws.launch_lsqml_group(g, constraints=self.group_constraints, **flags)


class EPIESolver(_SolverBase):
    """phaser/engines/conventional/solvers.py:293-383"""

    @classmethod
    def name(cls):
        return "ePIE"

    def _set_params(self, sim):
        sim.workspace.set_params(beta_object=self._schedule('beta_object', sim),
                                 beta_probe=self._schedule('beta_probe', sim))

    def _launch_group(self, ws, g, **flags):
        ws.launch_epie_group(g, constraints=self.group_constraints, **flags)


SOLVER_CLASSES = {'lsqml': LSQMLSolver, 'epie': EPIESolver}


def run_engine(args, props, *, use_graphs=True, keep_chi=None, fused=True):
    """Conventional engine (phaser/engines/conventional/run.py:15-122).

    args: EngineArgs mapping {data: Patterns, state: ReconsState, dtype, xp, recons_name,
    observer, seed}.  state.object.data / state.probe.data may be numpy arrays or torch
    tensors; the returned state holds torch tensors on the GPU.

    keep_chi (LSQ-ML; default on): keep every slice's back-propagated chi, (S-1) B M n^2
    complex64, so a group's object updates run as one launch after the backward sweep
    instead of one per slice inside it.  fused: whole-sweep kernels (one launch for the
    forward model, one for the LSQ-ML backward sweep) instead of one launch per slice."""
    dtype = args.get('dtype', numpy.float32)
    if numpy.dtype(dtype) != numpy.float32:
        raise NotImplementedError("the MI355X backend runs float32 / complex64 plans")
    observer = args.get('observer') or Observer()
    seed = args.get('seed')
    state = args['state']
    if not hasattr(state.object.data, 'is_cuda'):
        state = state.to_device()

    noise = props.noise_model
    if noise.name == 'poisson':
        raise NotImplementedError()  # PoissonNoiseModel.calc_wave_update (noise_models.py:104-111)

    update_probe = process_flag(props.update_probe)
    update_object = process_flag(props.update_object)
    update_positions = process_flag(props.update_positions)
    calc_error = process_flag(props.calc_error)
    shuffle_groups = shuffle_groups_flag(props)

    patterns = host_patterns(args['data'].patterns)   # numpy, or float32 already on the GPU
    pattern_mask = numpy.asarray(args['data'].pattern_mask)
    assert (hasattr(patterns, 'is_cuda') or patterns.dtype == numpy.float32) and pattern_mask.dtype == numpy.float32

    solver = SOLVER_CLASSES[props.solver.name](props, props.solver, use_graphs=use_graphs)
    groups = GroupManager(state.scan, props.grouping, props.compact, seed=seed)
    calc_error_mask = mask_fraction_of_groups(len(groups), props.calc_error_fraction)
    position_solver = None if props.position_solver is None else make_position_solver(props.position_solver)
    position_solver_state = None if position_solver is None else position_solver.init_state(state)

    observer.init_engine(state, recons_name=args.get('recons_name', ''), plan=props, noise_model=noise.name)
    start_i = state.iter.total_iter

    propagators = make_propagators(state.probe.sampling, state.wavelength, state.object.thicknesses, props.bwlim_frac)
    n_pos = int(numpy.prod(state.scan.shape[:-1]))
    bmax = int(numpy.ceil(n_pos / len(groups)))
    ws = GroupWorkspace(state, patterns, pattern_mask, propagators, noise_offset=noise.offset, noise_eps=noise.eps,
                        n_groups=len(groups), bmax=bmax,
                        keep_chi=(props.solver.name == 'lsqml') if keep_chi is None else bool(keep_chi),
                        fused=fused)
    ws.set_positions(state.object.sampling, state.scan)
    # constraint tables are built once here (simulation.py:127-132), outside any graph
    solver.group_constraints = tuple(make_constraints(props.group_constraints, ws, state, 'group constraint'))
    iter_constraints = make_constraints(props.iter_constraints, ws, state, 'iter constraint')
    sim = SimulationState(state=state, noise_model=noise, workspace=ws)
    sim = solver.init(sim)

    sim = solver.presolve(sim, groups.iter(state.scan), patterns=patterns, pattern_mask=pattern_mask,
                          propagators=propagators)
    observer.start_engine(sim.state)

    for i in range(1, props.niter + 1):
        state.iter.engine_iter = i
        state.iter.total_iter = start_i + i
        fargs = {'state': state, 'niter': props.niter}
        iter_update_positions = update_positions(fargs)
        iter_shuffle_groups = shuffle_groups(fargs)

        sim, pos_update, group_errors = solver.run_iteration(
            sim, groups.iter(state.scan, i, iter_shuffle_groups),
            patterns=patterns, pattern_mask=pattern_mask, propagators=propagators,
            update_object=update_object(fargs), update_probe=update_probe(fargs),
            update_positions=iter_update_positions, calc_error=calc_error(fargs),
            calc_error_mask=calc_error_mask, observer=observer)

        # iter constraints (run.py:93); a NaN they make is found by the next iteration's
        # state check, where the reference's next finite check finds it
        for con in iter_constraints:
            con.apply(ws)

        if iter_update_positions:
            if not position_solver:
                raise ValueError("Updating positions with no PositionSolver specified")
            axes = tuple(range(pos_update.ndim - 1))
            pos_update -= numpy.mean(pos_update, axes)
            pos_update, position_solver_state = position_solver.perform_update(state.scan, pos_update,
                                                                               position_solver_state)
            pos_update -= numpy.mean(pos_update, axes)
            logger.info(f"Position update: mean {numpy.mean(numpy.linalg.norm(pos_update, axis=-1))}")
            state.scan += pos_update
            state.object.sampling.check_scan(state.scan, state.probe.sampling.extent / 2.)
            ws.set_positions(state.object.sampling, state.scan)

        error = None
        if group_errors is not None and len(group_errors):
            error = float(numpy.nanmean(numpy.concatenate(group_errors)))
            state.progress.iters = numpy.concatenate([state.progress.iters, [i + start_i]])
            state.progress.detector_errors = numpy.concatenate([state.progress.detector_errors, [error]])
        observer.update_iteration(state, i, props.niter, error)

    observer.finish_engine(state)
    return state
