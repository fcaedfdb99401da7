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


def run_phaser_engine(args, plan_data, engine_type):
    """Run one engine on phaser's EngineArgs and plan mapping; returns phaser's state
    object, updated in place (the contract of phaser/hooks/__init__.py:174-185)."""
    if engine_type == 'conventional':
        from .engines.conventional import run_engine
    elif engine_type == 'gradient':
        from .engines.gradient import run_engine
    else:
        raise ValueError(f"unknown engine type {engine_type!r}")
    ours = dict(args)
    ours['state'] = state_from_phaser(args['state'])
    ours['data'] = patterns_from_phaser(args['data'])
    ours['observer'] = None   # phaser's observer expects its own state class
    out = run_engine(ours, plan_from_phaser(plan_data, engine_type))
    return state_into_phaser(out.to_numpy(), args['state'])
