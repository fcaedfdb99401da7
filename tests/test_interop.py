"""The phaser-side adapters (phaser_amd/interop.py) on objects with phaser's attribute
layout (phaser/state.py:17-211); phaser itself is not importable here."""
from types import SimpleNamespace as NS

import numpy

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
