"""Adapters between phaser's own state/plan objects and this backend's.

phaser is not imported: the adapters read and write the attributes phaser's classes
define (phaser/state.py:17-211, phaser/utils/num.py:544-670, phaser/utils/object.py:189-
481), so any object with that layout works.  Plans cross as the mapping form of a plan
YAML `engines:` entry, which is what `pane.into_data(props)` produces for phaser's
ConventionalEnginePlan / GradientEnginePlan (phaser/plan.py:43-173)."""
import numpy

from .plan import from_dict
from .sampling import ObjectSampling, Sampling
from .state import IterState, ObjectState, Patterns, ProbeState, ProgressState, ReconsState


def _np(a):
    if a is None:
        return None
    if hasattr(a, 'detach'):
        return a.detach().cpu().numpy()
    return numpy.asarray(a)


def state_from_phaser(st) -> ReconsState:
    """phaser ReconsState (numpy arrays) -> this backend's ReconsState (numpy arrays)."""
    ps, os_ = st.probe.sampling, st.object.sampling
    probe = ProbeState(Sampling(tuple(int(v) for v in ps.shape), sampling=tuple(float(v) for v in ps.sampling)),
                       numpy.ascontiguousarray(_np(st.probe.data)))
    osamp = ObjectSampling(tuple(int(v) for v in os_.shape), _np(os_.sampling), _np(os_.corner),
                           _np(getattr(os_, 'region_min', None)), _np(getattr(os_, 'region_max', None)))
    obj = ObjectState(osamp, numpy.ascontiguousarray(_np(st.object.data)), _np(st.object.thicknesses))
    it = st.iter
    iters = IterState(engine_num=int(it.engine_num), engine_iter=int(it.engine_iter), total_iter=int(it.total_iter),
                      n_engine_iters=it.n_engine_iters, n_total_iters=it.n_total_iters)
    prog = getattr(st, 'progress', None)
    progress = ProgressState() if prog is None else ProgressState(
        iters=numpy.array(_np(prog.iters)), detector_errors=numpy.array(_np(prog.detector_errors)))
    return ReconsState(iter=iters, wavelength=float(st.wavelength), probe=probe, object=obj,
                       scan=numpy.array(_np(st.scan)), tilt=_np(getattr(st, 'tilt', None)), progress=progress)


def state_into_phaser(ours: ReconsState, target):
    """Write a result of this backend back into phaser's ReconsState `target` (numpy
    arrays, dtypes kept) and return it."""
    target.probe.data = numpy.asarray(_np(ours.probe.data), dtype=numpy.asarray(target.probe.data).dtype)
    target.object.data = numpy.asarray(_np(ours.object.data), dtype=numpy.asarray(target.object.data).dtype)
    target.scan = numpy.asarray(_np(ours.scan), dtype=numpy.asarray(target.scan).dtype)
    for k in ('engine_num', 'engine_iter', 'total_iter', 'n_engine_iters', 'n_total_iters'):
        setattr(target.iter, k, getattr(ours.iter, k))
    if getattr(target, 'progress', None) is not None:
        target.progress.iters = numpy.asarray(ours.progress.iters)
        target.progress.detector_errors = numpy.asarray(ours.progress.detector_errors)
    return target


def patterns_from_phaser(p) -> Patterns:
    return Patterns(numpy.asarray(_np(p.patterns)), numpy.asarray(_np(p.pattern_mask)))


def plan_from_phaser(plan_data, engine_type):
    """Engine plan from the mapping form of phaser's plan (`pane.into_data(props)`)."""
    d = dict(plan_data)
    d['type'] = engine_type
    return from_dict(d)


def _overrides(obs, name):
    """Whether phaser observer `obs` does anything in callback `name`: a method of its
    own, not the no-op of phaser's base `Observer` (observer.py:18-69).  phaser's
    ObserverSet (observer.py:271-332) forwards to its children, so they decide."""
    children = getattr(obs, 'inner', None)
    if isinstance(children, (tuple, list)):
        return any(_overrides(c, name) for c in children)
    fn = getattr(type(obs), name, None)
    return fn is not None and getattr(fn, '__qualname__', '') != f'Observer.{name}'


class PhaserObserver:
    """The engines' observer callbacks forwarded to one of phaser's observers
    (LoggingObserver, PatienceObserver, SaveObserver, an ObserverSet of them; observer.py)
    with phaser's own state object: before each callback the engine's state is copied
    into `target` (state_into_phaser), so the observer sees the arrays, iteration counters
    and progress it would see on the numpy backend.  `plan` is phaser's engine plan object,
    handed to `init_engine` (SaveObserver reads its save flags and options there).

    An EarlyTermination raised by the observer (PatienceObserver, observer.py:142-181)
    propagates unchanged, carrying `target`, which is already up to date.  Per-group
    copies are made only when the observer implements `update_group`
    (`wants_groups`); the engines then call it after every group."""

    def __init__(self, inner, target, plan=None):
        self.inner = inner
        self.target = target
        self.plan = plan
        self.wants_groups = _overrides(inner, 'update_group')
        self._wants_iter = _overrides(inner, 'update_iteration')

    def _sync(self, state):
        return state_into_phaser(state.to_numpy() if hasattr(state, 'to_numpy') else state, self.target)

    def init_engine(self, state, *, recons_name, plan, **kwargs):
        self.inner.init_engine(self._sync(state), recons_name=recons_name,
                               plan=self.plan if self.plan is not None else plan, **kwargs)

    def start_engine(self, state):
        self.inner.start_engine(self._sync(state))

    def heartbeat(self):
        if hasattr(self.inner, 'heartbeat'):
            self.inner.heartbeat()

    def update_group(self, state, force=False):
        if self.wants_groups:
            self.inner.update_group(self._sync(state), force)

    def update_iteration(self, state, i, n, error=None):
        self.inner.update_iteration(self._sync(state) if self._wants_iter else self.target, i, n, error)

    def finish_engine(self, state):
        self.inner.finish_engine(self._sync(state))


def run_phaser_engine(args, plan_data, engine_type, props=None):
    """Run one engine on phaser's EngineArgs and plan mapping; returns phaser's state
    object, updated in place (the contract of phaser/hooks/__init__.py:174-185).

    phaser's observer (`args['observer']`) receives the engine's callbacks through
    PhaserObserver; `props` is phaser's plan object, given to its `init_engine`."""
    if engine_type == 'conventional':
        from .engines.conventional import run_engine
    elif engine_type == 'gradient':
        from .engines.gradient import run_engine
    else:
        raise ValueError(f"unknown engine type {engine_type!r}")
    ours = dict(args)
    ours['state'] = state_from_phaser(args['state'])
    ours['data'] = patterns_from_phaser(args['data'])
    obs = args.get('observer')
    ours['observer'] = None if obs is None else PhaserObserver(obs, args['state'], props)
    out = run_engine(ours, plan_from_phaser(plan_data, engine_type))
    return state_into_phaser(out.to_numpy(), args['state'])
