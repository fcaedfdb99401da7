"""The phaser-side adapters (phaser_amd/interop.py) on objects with phaser's attribute
layout (phaser/state.py:17-211); phaser itself is not importable here."""
from types import SimpleNamespace as NS

import numpy
import pytest

from phaser_amd.interop import patterns_from_phaser, plan_from_phaser, state_from_phaser, state_into_phaser
from phaser_amd.plan import ConventionalEnginePlan, GradientEnginePlan


def _phaser_like_state():
    rng = numpy.random.default_rng(0)
    probe = NS(sampling=NS(shape=numpy.array([16, 16]), sampling=numpy.array([0.2, 0.2])),
               data=(rng.normal(size=(2, 16, 16)) + 1j).astype(numpy.complex64))
    osamp = NS(shape=numpy.array([40, 44]), sampling=numpy.array([0.2, 0.2]), corner=numpy.array([-3.9, -4.3]),
               region_min=numpy.array([-1.0, -1.0]), region_max=numpy.array([1.0, 1.0]))
    obj = NS(sampling=osamp, data=numpy.ones((3, 40, 44), numpy.complex64), thicknesses=numpy.array([20., 20., 20.]))
    it = NS(engine_num=1, engine_iter=0, total_iter=5, n_engine_iters=10, n_total_iters=20)
    prog = NS(iters=numpy.array([1, 2], numpy.uint64), detector_errors=numpy.array([3.0, 2.0]))
    return NS(iter=it, wavelength=0.0197, probe=probe, object=obj, scan=numpy.zeros((4, 4, 2), numpy.float32),
              tilt=None, progress=prog)


def test_state_round_trip():
    st = _phaser_like_state()
    ours = state_from_phaser(st)
    assert tuple(ours.object.sampling.shape) == (40, 44)
    numpy.testing.assert_array_equal(ours.object.sampling.corner, [-3.9, -4.3])
    numpy.testing.assert_array_equal(ours.probe.data, st.probe.data)
    assert ours.iter.total_iter == 5 and ours.wavelength == 0.0197
    ours.object.data = ours.object.data * 2
    ours.iter.total_iter = 15
    ours.progress.detector_errors = numpy.array([3.0, 2.0, 1.0])
    out = state_into_phaser(ours, st)
    assert out is st and st.object.data.dtype == numpy.complex64
    numpy.testing.assert_array_equal(st.object.data, 2)
    assert st.iter.total_iter == 15 and len(st.progress.detector_errors) == 3


def test_plan_and_patterns_from_mapping_form():
    p = plan_from_phaser({'niter': 7, 'solver': 'lsqml', 'noise_model': 'amplitude', 'grouping': 32,
                          'group_constraints': [], 'iter_constraints': []}, 'conventional')
    assert isinstance(p, ConventionalEnginePlan) and p.niter == 7 and p.grouping == 32
    g = plan_from_phaser({'niter': 3, 'noise_model': 'poisson', 'solvers': {}}, 'gradient')
    assert isinstance(g, GradientEnginePlan) and g.niter == 3
    pats = patterns_from_phaser(NS(patterns=numpy.zeros((2, 8, 8), numpy.float32),
                                   pattern_mask=numpy.ones((8, 8), numpy.float32)))
    assert pats.patterns.shape == (2, 8, 8)


# --- observer forwarding (phaser/observer.py:18-69) --------------------------------------
class Observer:
    """Stand-in for phaser's base observer: every callback a no-op (observer.py:18-69)."""

    def init_engine(self, init_state, *, recons_name, plan, **kwargs): pass
    def start_engine(self, init_state): pass
    def heartbeat(self): pass
    def update_group(self, state, force=False): pass
    def update_iteration(self, state, i, n, error=None): pass
    def finish_engine(self, state): pass


class Recorder(Observer):
    """Records what it sees, as LoggingObserver / SaveObserver would read it."""

    def __init__(self):
        self.calls = []

    def init_engine(self, init_state, *, recons_name, plan, **kwargs):
        self.calls.append(('init_engine', init_state.iter.total_iter, recons_name, plan, kwargs))

    def update_iteration(self, state, i, n, error=None):
        self.calls.append(('update_iteration', float(numpy.abs(state.object.data).max()), i, n, error))

    def finish_engine(self, state):
        self.calls.append(('finish_engine', state.iter.engine_iter))


class EarlyTermination(BaseException):
    """Stand-in for phaser.types.EarlyTermination (types.py:85-88)."""

    def __init__(self, state, continue_next_engine=False):
        self.state = state
        self.continue_next_engine = continue_next_engine


class Patience(Observer):
    """PatienceObserver's contract (observer.py:142-181): raise after `patience`
    iterations without improvement of state.progress.detector_errors[-1]."""

    def __init__(self, patience):
        self.patience, self.best, self.bad = patience, None, 0

    def update_iteration(self, state, i, n, error=None):
        e = state.progress.detector_errors[-1]
        if self.best is None or e < self.best:
            self.best, self.bad = e, 0
        else:
            self.bad += 1
        if self.bad >= self.patience:
            raise EarlyTermination(state, True)


def test_phaser_observer_sees_phaser_state_after_each_callback():
    from phaser_amd.interop import PhaserObserver
    target = _phaser_like_state()
    ours = state_from_phaser(target)
    rec = Recorder()
    phaser_plan = NS(niter=2, save=False)
    obs = PhaserObserver(rec, target, plan=phaser_plan)
    assert not obs.wants_groups            # Recorder keeps the base update_group
    obs.init_engine(ours, recons_name='r', plan='ours', noise_model='amplitude')
    ours.object.data = ours.object.data * 3
    ours.iter.engine_iter = 1
    obs.update_group(ours)                 # a no-op for this observer: no copy made
    assert numpy.abs(target.object.data).max() == 1
    obs.update_iteration(ours, 1, 2, 0.5)
    obs.finish_engine(ours)
    assert rec.calls[0] == ('init_engine', 5, 'r', phaser_plan, {'noise_model': 'amplitude'})
    assert rec.calls[1] == ('update_iteration', 3.0, 1, 2, 0.5)
    assert rec.calls[2] == ('finish_engine', 1)
    assert numpy.abs(target.object.data).max() == 3


def test_phaser_observer_group_updates_and_observer_sets():
    from phaser_amd.engines.conventional import wants_group_updates
    from phaser_amd.interop import PhaserObserver

    class PerGroup(Observer):
        def __init__(self):
            self.seen = []

        def update_group(self, state, force=False):
            self.seen.append((float(numpy.abs(state.object.data).max()), force))

    class ObserverSet(Observer):      # phaser's ObserverSet forwards to `inner`
        def __init__(self, inner):
            self.inner = tuple(inner)

        def update_group(self, state, force=False):
            for o in self.inner:
                o.update_group(state, force)

    target = _phaser_like_state()
    ours = state_from_phaser(target)
    pg = PerGroup()
    assert not PhaserObserver(ObserverSet([Recorder()]), target).wants_groups
    obs = PhaserObserver(ObserverSet([Recorder(), pg]), target)
    assert obs.wants_groups and wants_group_updates(obs)
    ours.object.data = ours.object.data * 2
    obs.update_group(ours, True)
    assert pg.seen == [(2.0, True)]
    # this backend's own no-op observer asks for no per-group callbacks
    from phaser_amd.engines.conventional import Observer as OwnObserver
    assert not wants_group_updates(OwnObserver())

    class Mine(OwnObserver):
        def update_group(self, state, force=False):
            pass
    assert wants_group_updates(Mine())


def test_phaser_observer_early_termination_carries_phaser_state():
    import pytest
    from phaser_amd.interop import PhaserObserver
    target = _phaser_like_state()
    ours = state_from_phaser(target)
    obs = PhaserObserver(Patience(1), target)
    ours.progress.detector_errors = numpy.array([3.0])
    obs.update_iteration(ours, 1, 5, 3.0)
    ours.progress.detector_errors = numpy.array([3.0, 4.0])
    with pytest.raises(EarlyTermination) as ei:
        obs.update_iteration(ours, 2, 5, 4.0)
    assert ei.value.state is target and list(target.progress.detector_errors) == [3.0, 4.0]


# --- group shuffling and position solvers at the boundary --------------------------------
def test_shuffle_groups_follows_reference_resolution():
    """`props.shuffle_groups or not props.compact` (conventional/run.py:32): an explicit
    False still shuffles sparse groups."""
    from phaser_amd.plan import Flag, shuffle_groups_flag
    args = {'state': NS(iter=NS(engine_iter=3)), 'niter': 5}
    for shuffle, compact, want in ((None, False, True), (False, False, True), (True, False, True),
                                   (None, True, False), (False, True, False), (True, True, True)):
        p = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', shuffle_groups=shuffle, compact=compact)
        assert shuffle_groups_flag(p)(args) is want, (shuffle, compact)
    p = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', shuffle_groups=Flag(after=4), compact=True)
    assert shuffle_groups_flag(p)(args) is False


def test_oracle_shuffles_sparse_groups_with_shuffle_false():
    """The oracle engine loop resolves the flag the same way: with compact=False and
    shuffle_groups=False the group order changes between iterations."""
    import oracle.misc as omisc
    scan = numpy.stack(numpy.meshgrid(numpy.arange(6.), numpy.arange(6.), indexing='ij'), axis=-1)
    gm = omisc.GroupManager(scan, 6, False, seed=4)
    flag = False or not False
    orders = [tuple(map(tuple, numpy.asarray(g).reshape(2, -1).T.tolist()) for g in gm.iter(scan, i, flag))
              for i in (1, 2)]
    assert orders[0] != orders[1]


def test_position_solvers_take_the_reference_protocol():
    """cls(args, props), init_state(sim) on the ReconsState (hooks/solver.py:20-75)."""
    from phaser_amd.engines.position import make_position_solver, MomentumPositionSolver
    from phaser_amd.plan import MomentumPositionSolverPlan, SteepestDescentPositionSolverPlan
    sim = NS(scan=numpy.ones((3, 4, 2), numpy.float32))
    mom = make_position_solver(MomentumPositionSolverPlan(step_size=0.5, momentum=0.5))
    assert isinstance(mom, MomentumPositionSolver)
    st = mom.init_state(sim)
    assert st.shape == (3, 4, 2) and not st.any()
    g = numpy.full((3, 4, 2), 2.0, numpy.float32)
    u, st = mom.perform_update(sim.scan, g, st)
    numpy.testing.assert_allclose(u, 1.0)
    u, st = mom.perform_update(sim.scan, g, st)
    numpy.testing.assert_allclose(u, 1.5)
    sd = make_position_solver(SteepestDescentPositionSolverPlan(step_size=0.1))
    assert sd.init_state(sim) is None and sd.max_step_size == 0.1   # position_correction.py:16


def test_tilt_tables():
    """Tilted propagators' host tables: nothing to tilt for no tilts, all-zero tilts or a
    single slice; otherwise tan(1e-3 tilt) per position in scan order (y, x) and the
    thicknesses of the S-1 propagators; a tilt map of the wrong size is an error."""
    from phaser_amd.device import tilt_tables
    th = numpy.array([20., 30., 20.])
    assert tilt_tables(None, th, 9) == (None, None)
    z = numpy.zeros((3, 3, 2), dtype=numpy.float32)
    assert tilt_tables(z, th, 9) == (None, None)
    t = z.copy()
    t[1, 2] = (5.0, -2.0)
    assert tilt_tables(t, th[:1], 9) == (None, None)
    tan, dz = tilt_tables(t, th, 9)
    assert tan.dtype == numpy.float32 and tan.shape == (9, 2) and dz.dtype == numpy.float32
    assert tan[5, 0] == numpy.float32(numpy.tan(5e-3)) and tan[5, 1] == numpy.float32(numpy.tan(-2e-3))
    assert not numpy.any(numpy.delete(tan, 5, axis=0))
    assert list(dz) == [20., 30.]
    with pytest.raises(ValueError, match="positions"):
        tilt_tables(t, th, 8)
