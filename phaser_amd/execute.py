"""Executing a reconstruction plan (reference phaser/execute.py:19-402, observers
phaser/observer.py:72-258): load the raw data and run the post_load hooks, initialise
probe, scan, tilt and object, run the post_init hooks, then each engine in turn after
preparing the state for it (object padding and resampling, probe modes, re-slicing,
pattern resampling to `sim_shape`).  The engines are this package's HIP engines whatever
`backend` the plan names: the MI355X backend replaces numpy/cupy/jax for every engine,
the gradient engine included.

Everything here runs once per engine on the host; the per-group hot path is the engines'.
"""
import itertools
import logging
import math
import time
from pathlib import Path

import numpy

from . import hooks
from .engines.conventional import Observer as _EngineObserver
from .plan import GradientEnginePlan, Flag, process_flag
from .sampling import ObjectSampling, Sampling
from .state import EarlyTermination, IterState, Patterns, ProgressState, ReconsState

logger = logging.getLogger(__name__)


# ------------------------------------------------------------------------ observers
class Observer(_EngineObserver):
    """Reconstruction lifecycle callbacks (observer.py:18-69); every one a no-op.  The
    engine callbacks are inherited from the engines' own no-op observer, so subclasses
    that leave `update_group` alone do not ask the engines for per-group state."""

    def init_recons(self, plan): pass
    def start_recons(self, init_state): pass
    def heartbeat(self): pass
    def finish_recons(self, state): pass
    def close(self, exc=None): pass


class ObserverSet(Observer):
    """Forwards every callback to each observer in turn (observer.py:271-332)."""

    def __init__(self, observers):
        self.inner = tuple(observers)

    @property
    def wants_groups(self):
        from .engines.conventional import wants_group_updates
        return any(wants_group_updates(o) for o in self.inner)

    def _each(name):
        def f(self, *a, **k):
            for o in self.inner:
                getattr(o, name)(*a, **k)
        f.__name__ = name
        return f

    init_recons = _each('init_recons')
    start_recons = _each('start_recons')
    init_engine = _each('init_engine')
    start_engine = _each('start_engine')
    heartbeat = _each('heartbeat')
    update_group = _each('update_group')
    update_iteration = _each('update_iteration')
    finish_engine = _each('finish_engine')
    finish_recons = _each('finish_recons')
    close = _each('close')
    del _each


class LoggingObserver(Observer):
    """Progress and timing through `logging` (observer.py:72-140)."""

    def __init__(self):
        self.t_init = self.t_recons = self.t_engine = self.t_iter = None

    def init_recons(self, plan):
        logger.info("Initializing reconstruction...")
        self.t_init = time.monotonic()

    def start_recons(self, init_state):
        self.t_recons = time.monotonic()
        logger.info("Initialized reconstruction" + (
            f" in {self.t_recons - self.t_init:.3f} s" if self.t_init is not None else ""))

    def init_engine(self, init_state, *, recons_name, plan, **kwargs):
        logger.info("Initializing engine...")
        self.t_engine = time.monotonic()

    def start_engine(self, init_state):
        logger.info("Engine initialized")
        self.t_iter = time.monotonic()

    def update_iteration(self, state, i, n, error=None):
        t = time.monotonic()
        dt = f" [{t - self.t_iter:.3f} s]" if self.t_iter is not None else ""
        w = len(str(n))
        err = f" Error: {error:.3e}" if error is not None else ""
        logger.info(f"Finished iter {i:{w}}/{n}{dt}{err}")
        self.t_iter = t

    def finish_engine(self, state):
        logger.info("Engine finished!" + (
            f" Total engine time: {time.monotonic() - self.t_engine:.3f} s" if self.t_engine is not None else ""))

    def finish_recons(self, state):
        logger.info("Finished reconstruction!")


class PatienceObserver(Observer):
    """Early termination after `patience` iterations without a new best detector error
    (observer.py:142-181)."""

    def __init__(self, patience, smoothing=0.1, continue_next_engine=True):
        self.patience, self.smoothing, self.continue_next_engine = patience, smoothing, continue_next_engine
        self.no_improvement_iter = 0
        self.best_error = None
        self.smoothed_error = None

    def init_engine(self, init_state, *, recons_name, plan, **kwargs):
        self.no_improvement_iter = 0

    def update_iteration(self, state, i, n, error=None):
        errs = getattr(state.progress, 'detector_errors', None)
        if errs is None or numpy.size(errs) == 0:
            return
        error = errs[-1]
        if self.best_error is None or error < self.best_error:
            self.best_error = error
            self.no_improvement_iter = 0
        else:
            self.no_improvement_iter += 1
        self.smoothed_error = error if self.smoothed_error is None else \
            (1 - self.smoothing) * self.smoothed_error + self.smoothing * error
        if self.no_improvement_iter >= self.patience:
            logger.info(f"Early termination: no improvement for {self.patience} iterations")
            raise EarlyTermination(state, self.continue_next_engine)


def flag_any_true(flag, niter):
    """Whether a flag fires in any of an engine's iterations (types.py:169-180)."""
    if isinstance(flag, bool):
        return flag
    if isinstance(flag, Flag):
        return flag.after < niter and (flag.before is None or flag.before > 0)
    return True


class SaveObserver(Observer):
    """Writes the reconstruction state on the plan's `save` flag and at the end of each
    engine (observer.py:183-258), in the checkpoint layout of phaser/utils/io.py:176-245
    (phaser_amd.io).  Images (`save_images`) need phaser's matplotlib/tifffile writers and
    are not produced; the request is logged."""

    def __init__(self):
        self.out_dir = None

    def init_engine(self, init_state, *, recons_name, plan, **kwargs):
        from .plan import SaveOptions
        self.plan = plan
        self.opts = plan.save_options or SaveOptions()
        self.save_flag = process_flag(plan.save)
        self.any_state = flag_any_true(plan.save, plan.niter)
        if flag_any_true(plan.save_images, plan.niter):
            logger.warning("save_images: image output is not produced by the MI355X front-end")
        try:
            out = self.opts.out_dir.format(engine_num=init_state.iter.engine_num, name=recons_name,
                                           group=plan.grouping, niter=plan.niter, **kwargs)
        except KeyError as e:
            raise ValueError(f"Invalid format string in 'out_dir' (unknown key {e})") from None
        self.out_dir = Path(out).expanduser().absolute()
        if self.any_state:
            self.out_dir.mkdir(exist_ok=True, parents=True)
            (self.out_dir / 'finished').unlink(missing_ok=True)

    def _save(self, state):
        from .io import write_state
        path = self.out_dir / self.opts.hdf5_fmt.format(iter=state.iter, type='state')
        if path.suffix.lower() in ('.h5', '.hdf5'):
            try:
                import h5py  # noqa: F401
            except ImportError:
                path = path.with_suffix('.npz')   # same layout, numpy container
        write_state(state, path)

    def update_iteration(self, state, i, n, error=None):
        if self.any_state and self.save_flag({'state': state, 'niter': n}):
            self._save(state)

    def finish_engine(self, state):
        if self.any_state:
            self._save(state)

    def close(self, exc=None):
        if exc is None and self.out_dir is not None and self.any_state:
            (self.out_dir / 'finished').touch(mode=0o664)


def _normalize_observers(observers, override_observers):
    """execute.py:94-122."""
    if override_observers is not None:
        if observers is not None:
            raise TypeError("Cannot specify both 'observers' and 'override_observers'")
        obs = (override_observers,) if isinstance(override_observers, Observer) else tuple(override_observers)
        return ObserverSet(obs)
    obs = [SaveObserver(), LoggingObserver()]
    if isinstance(observers, Observer):
        obs.append(observers)
    elif observers is not None:
        obs.extend(observers)
    return ObserverSet(obs)


# ------------------------------------------------------------------------ plan steps
class PreparedRecons:
    def __init__(self, patterns, state, name, observer):
        self.patterns, self.state, self.name, self.observer = patterns, state, name, observer


def load_raw_data(plan, seed=None, init_state=None):
    """execute.py:125-187: the raw_data hook, wavelength, init overrides of its probe /
    scan / tilt hooks, dtype, then the post_load hooks."""
    dtype = numpy.float32 if plan.dtype == 'float32' else numpy.float64
    raw = hooks.resolve(plan.raw_data, hooks.RAW_DATA, 'raw_data')(None)
    wavelength = plan.wavelength or raw.get('wavelength')
    if wavelength is None:
        raise ValueError("`wavelength` must be specified by raw_data or manually")
    for key, hook in (('scan', 'scan_hook'), ('tilt', 'tilt_hook'), ('probe', 'probe_hook')):
        manual = plan.init.get(key)
        if manual not in (None, {}):
            base = dict(raw.get(hook) or {})
            man = {'type': manual} if isinstance(manual, str) else dict(manual)
            if base.get('type') not in (None, man.get('type')):
                base = {}
            base.update(man)
            raw[hook] = base
        elif manual == {} and raw.get(hook) is None:
            raise ValueError(f"Manual `init.{key}` specified to override initial state, but {key} was not "
                             "provided by the raw data")
    if raw.get('scan_hook') is None and (init_state is None or init_state.scan is None):
        raise ValueError("`scan` must be specified by raw data, previous state, or manually in `init.scan`")
    if raw.get('probe_hook') is None and (init_state is None or init_state.probe is None):
        raise ValueError("`probe` must be specified by raw data, previous state, or manually in `init.probe`")
    raw['wavelength'] = wavelength
    raw['seed'] = seed
    if raw['patterns'].dtype != dtype:
        raw['patterns'] = raw['patterns'].astype(dtype)
    for p in plan.post_load:
        raw = hooks.resolve(p, hooks.POST_LOAD, 'post_load')(raw)
    if isinstance(raw['patterns'], numpy.memmap):
        raw['patterns'] = numpy.array(raw['patterns'])
    return raw


def initialize_reconstruction(plan, *, seed=None, name=None, init_state=None, observers=None,
                              override_observers=None):
    """execute.py:190-331."""
    observer = _normalize_observers(observers, override_observers)
    observer.init_recons(plan)
    dtype = numpy.float32 if plan.dtype == 'float32' else numpy.float64
    cdtype = numpy.complex64 if dtype == numpy.float32 else numpy.complex128
    if init_state is None and plan.init.get('state') is not None:
        from .io import read_state
        init_state = read_state(Path(plan.init['state']).expanduser())
    raw = load_raw_data(plan, seed, init_state)
    data = Patterns(raw['patterns'], raw['mask'])
    sampling = raw['sampling']
    wavelength = raw['wavelength']
    args = {'sampling': sampling, 'wavelength': wavelength, 'dtype': dtype, 'seed': seed}
    if init_state is not None and init_state.probe is not None and plan.init.get('probe') is None:
        probe = init_state.probe
        probe.data = numpy.asarray(probe.data).astype(cdtype)
        if probe.sampling != sampling:
            data.patterns = sampling.resample_recip(data.patterns, probe.sampling)
            data.pattern_mask = sampling.resample_recip(data.pattern_mask, probe.sampling)
            sampling = probe.sampling
    else:
        probe = hooks.resolve(raw['probe_hook'], hooks.PROBE, 'probe')(args)
    if probe.data.ndim == 2:
        probe.data = probe.data.reshape((1, *probe.data.shape))
    if init_state is not None and init_state.scan is not None and plan.init.get('scan') is None:
        scan = init_state.scan
    else:
        scan = hooks.resolve(raw['scan_hook'], hooks.SCAN, 'scan')({'dtype': dtype, 'seed': seed})
    if init_state is not None and init_state.tilt is not None and plan.init.get('tilt') is None:
        tilt = init_state.tilt
    elif raw.get('tilt_hook') is not None:
        tilt = hooks.resolve(raw['tilt_hook'], hooks.TILT, 'tilt')({'dtype': dtype, 'shape': scan.shape[:-1]})
    else:
        tilt = None
    obj_pad_px = plan.engines[0].obj_pad_px if plan.engines else 5.0
    obj_sampling = ObjectSampling.from_scan(scan, sampling.sampling, sampling.extent / 2. + obj_pad_px * sampling.sampling)
    if init_state is not None and init_state.object is not None and plan.init.get('object') is None:
        obj = init_state.object
        obj.data = numpy.asarray(obj.data).astype(cdtype)
    else:
        obj = hooks.resolve(plan.init.get('object') or 'random', hooks.OBJECT, 'object')(
            {'sampling': obj_sampling, 'slices': plan.slices, 'wavelength': wavelength, 'dtype': dtype, 'seed': seed})
    if obj.data.ndim == 2:
        obj.data = obj.data.reshape((1, *obj.data.shape))
        obj.thicknesses = numpy.array([], dtype=dtype)
    state = ReconsState(iter=IterState(0, 0, 0), wavelength=wavelength, probe=probe, object=obj, scan=scan,
                        tilt=tilt, progress=ProgressState())
    for p in plan.post_init:
        data, state = hooks.resolve(p, hooks.POST_INIT, 'post_init')(
            {'data': data, 'state': state, 'dtype': dtype, 'seed': seed})
    if state.scan.shape[:-1] != data.patterns.shape[:-2]:
        n_pos, n_pat = math.prod(state.scan.shape[:-1]), math.prod(data.patterns.shape[:-2])
        if n_pos != n_pat:
            raise ValueError(f"# of scan positions {n_pos} doesn't match # of patterns {n_pat}")
        data.patterns = data.patterns.reshape((*state.scan.shape[:-1], *data.patterns.shape[-2:]))
    avg = float(numpy.nanmean(numpy.nansum(data.patterns, axis=(-1, -2))))
    if avg < 5.0:
        logger.warning(f"Mean pattern intensity is very low ({avg} particles). Ensure that it is being scaled "
                       "correctly to units of electrons/photons. For simulated data, use the 'scale' or 'poisson' "
                       "preprocessing")
    observer.start_recons(state)
    return PreparedRecons(data, state, name or plan.name, observer)


def prepare_for_engine(patterns, state, engine):
    """execute.py:334-402 on host arrays: pattern/probe resampling to `sim_shape`, object
    resampling to the probe's pixel size and padding to the scan, probe modes, slices."""
    state = state.to_numpy()
    if engine.sim_shape is not None and tuple(engine.sim_shape) != tuple(state.probe.data.shape[-2:]):
        if engine.resize_method == 'pad_crop':
            new = Sampling(engine.sim_shape, extent=tuple(state.probe.sampling.extent))
        else:
            new = Sampling(engine.sim_shape, sampling=tuple(state.probe.sampling.sampling))
        logger.info(f"Resampling probe and patterns to shape {tuple(new.shape)}...")
        state.probe.data = state.probe.sampling.resample(state.probe.data, new)
        patterns.patterns = state.probe.sampling.resample_recip(patterns.patterns, new)
        patterns.pattern_mask = state.probe.sampling.resample_recip(patterns.pattern_mask, new)
        state.probe.sampling = new
    osamp = state.object.sampling
    if not numpy.allclose(state.probe.sampling.sampling, osamp.sampling):
        osamp = osamp.with_sampling(state.probe.sampling.sampling)
    pad = osamp.expand_to_scan(state.scan, state.probe.sampling.extent / 2. + engine.obj_pad_px *
                               state.probe.sampling.sampling)
    if pad != osamp:
        osamp = pad
    if osamp != state.object.sampling:
        state.object.data = state.object.sampling.resample(state.object.data, osamp)
        state.object.sampling = osamp
    current = state.probe.data.shape[0]
    if engine.probe_modes != current:
        if engine.probe_modes < current:
            state.probe.data = state.probe.data[:engine.probe_modes]
        else:
            from .optics import make_hermetian_modes
            base = numpy.sum(state.probe.data, axis=0)
            state.probe.data = make_hermetian_modes(base, engine.probe_modes, base_mode_power=engine.base_mode_power)
    if engine.slices is not None:
        new_t = numpy.asarray(engine.slices, dtype=numpy.float64)
        old_t = numpy.asarray(state.object.thicknesses)
        if len(new_t) != len(old_t) or not numpy.allclose(new_t, old_t):
            from .optics import resample_slices
            state.object.data = resample_slices(state.object.data, old_t, new_t)
            state.object.thicknesses = numpy.array(new_t, dtype=old_t.dtype if old_t.size else numpy.float32)
    if isinstance(engine, GradientEnginePlan):
        solver_vars = set(itertools.chain.from_iterable(engine.solvers.keys()))
        if 'tilt' in solver_vars and state.tilt is None:
            state.tilt = numpy.zeros_like(state.scan)
    state.object.data = numpy.ascontiguousarray(state.object.data, dtype=numpy.complex64)
    state.probe.data = numpy.ascontiguousarray(state.probe.data, dtype=numpy.complex64)
    return patterns, state


def execute_engine(recons, engine, seed=None):
    """execute.py:47-91: one engine of the plan on the MI355X engines."""
    from .engines import conventional, gradient
    run = gradient.run_engine if isinstance(engine, GradientEnginePlan) else conventional.run_engine
    engine_i = recons.state.iter.engine_num
    obs = recons.observer
    if engine.early_termination:
        obs = ObserverSet((recons.observer, PatienceObserver(engine.early_termination,
                                                             engine.early_termination_smoothing)))
    logger.info(f"Preparing for engine #{engine_i + 1}...")
    recons.patterns, recons.state = prepare_for_engine(recons.patterns, recons.state, engine)
    recons.state.iter = IterState(engine_num=engine_i + 1, engine_iter=0, n_engine_iters=engine.niter,
                                  total_iter=recons.state.iter.total_iter,
                                  n_total_iters=recons.state.iter.n_total_iters)
    pats = recons.patterns
    if pats.patterns.dtype != numpy.float32:
        raise NotImplementedError("the MI355X backend runs float32 / complex64 plans")
    try:
        out = run({'data': pats, 'state': recons.state, 'dtype': pats.patterns.dtype, 'recons_name': recons.name,
                   'observer': obs, 'seed': seed}, engine)
        recons.state = out.to_numpy()
    except EarlyTermination as e:
        recons.state = e.state.to_numpy() if hasattr(e.state, 'to_numpy') else e.state
        obs.finish_engine(recons.state)
        if not e.continue_next_engine:
            obs.finish_recons(recons.state)
            raise
    return recons


def execute_plan(plan, *, seed=None, name=None, init_state=None, observers=None, override_observers=None):
    """execute.py:19-44: every engine of `plan` in turn; returns the final state.

    `seed` seeds the raw-data noise, the random object and every engine's groupings (the
    reference passes seed None to its engines, execute.py:81, so its group order is not
    reproducible; here one seed makes the whole plan reproducible)."""
    recons = initialize_reconstruction(plan, seed=seed, name=name, init_state=init_state, observers=observers,
                                       override_observers=override_observers)
    recons.state.iter.n_total_iters = sum(e.niter for e in plan.engines)
    try:
        try:
            for engine in plan.engines:
                recons = execute_engine(recons, engine, seed)
        except EarlyTermination as e:
            recons.state = e.state.to_numpy() if hasattr(e.state, 'to_numpy') else e.state
        recons.observer.finish_recons(recons.state)
        logger.info("Reconstruction finished!")
        recons.observer.close()
    except BaseException as exc:
        recons.observer.close(exc)
        raise
    return recons.state
