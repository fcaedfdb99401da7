"""The plan front-end on the CPU: YAML plans, hooks (raw data, preprocessing, initial
probe/scan/object), preparing a state for an engine, checkpoints and observers
(phaser_amd/plan.py, hooks.py, execute.py, io.py; reference phaser/plan.py, hooks/,
execute.py, utils/io.py, observer.py)."""
import os
from types import SimpleNamespace as NS

import numpy
import pytest

from phaser_amd import hooks
from phaser_amd.plan import ReconsPlan, ConventionalEnginePlan, GradientEnginePlan, plans_from_yaml
from phaser_amd.sampling import ObjectSampling, Sampling
from phaser_amd.state import IterState, ObjectState, Patterns, ProbeState, ProgressState, ReconsState

EXAMPLES = '/root/reference/examples'


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference example plans not present")
def test_reference_example_plans_parse():
    """Every example plan of the reference parses with its engines, solvers, constraints
    and regularizers resolved (their raw data files are not part of the reference)."""
    names = sorted(f for f in os.listdir(EXAMPLES) if f.endswith('.yaml'))
    assert names
    for f in names:
        for p in plans_from_yaml(os.path.join(EXAMPLES, f)):
            assert p.engines and all(isinstance(e, (ConventionalEnginePlan, GradientEnginePlan)) for e in p.engines)


PLAN = """
name: "synthetic_small"
backend: jax
raw_data:
  type: synthetic
  n: 32
  n_modes: 1
  n_slices: 2
  scan_shape: [6, 6]
  scan_step_px: 3.0
post_load:
 - type: poisson
slices:
  n: 2
  total_thickness: 40
engines:
 - type: conventional
   probe_modes: 2
   niter: 3
   grouping: 12
   noise_model: amplitude
   solver: {type: lsqml, beta_object: 0.5}
   group_constraints:
    - type: clamp_object_amplitude
      amplitude: 1.1
   iter_constraints:
    - {type: limit_probe_support, max_angle: 22.0}
   update_probe: {after: 1}
   save: {every: 2}
 - type: gradient
   niter: 2
   grouping: 12
   noise_model: {type: amplitude, eps: 1.0e-3}
   solvers:
     object: {type: sgd, learning_rate: 1.0, momentum: 0.5}
   regularizers:
    - {type: obj_l2, cost: 0.1}
   group_constraints: []
   iter_constraints: []
"""


def test_plan_yaml_fields(tmp_path):
    f = tmp_path / 'p.yaml'
    f.write_text(PLAN + "---\n" + PLAN.replace('synthetic_small', 'second'))
    plans = plans_from_yaml(f)
    assert [p.name for p in plans] == ['synthetic_small', 'second']
    p = plans[0]
    assert p.slices == [20.0, 20.0] and p.raw_data['type'] == 'synthetic'
    c, g = p.engines
    assert c.probe_modes == 2 and c.solver.beta_object == 0.5 and c.group_constraints[0].amplitude == 1.1
    assert c.update_probe({'state': NS(iter=NS(engine_iter=2)), 'niter': 3})
    assert not c.update_probe({'state': NS(iter=NS(engine_iter=1)), 'niter': 3})
    assert g.regularizers[0].name == 'obj_l2' and g.regularizers[0].cost == 0.1
    with pytest.raises(ValueError, match='version'):
        ReconsPlan(name='x', raw_data={'type': 'npz', 'path': 'x'}, version='2.0')
    with pytest.raises(ValueError, match='Unknown group constraint'):
        ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', group_constraints=[{'type': 'layers'}])


def test_raster_scan_and_focused_probe_hooks():
    scan = hooks.raster_scan({'dtype': numpy.float32}, {'shape': (3, 4), 'step_size': (2.0, 1.5)})
    assert scan.shape == (3, 4, 2) and scan.dtype == numpy.float32
    numpy.testing.assert_allclose(scan[:, 0, 0], [-3.0, -1.0, 1.0])          # (i - n/2) step
    numpy.testing.assert_allclose(scan[0, :, 1], [-3.0, -1.5, 0.0, 1.5])
    rot = hooks.raster_scan({'dtype': numpy.float64}, {'shape': (3, 4), 'step_size': 1.0, 'rotation': 90.0})
    numpy.testing.assert_allclose(numpy.linalg.norm(rot, axis=-1),
                                  numpy.linalg.norm(hooks.make_raster_scan((3, 4), 1.0), axis=-1))
    samp = Sampling((32, 32), sampling=(0.2, 0.2))
    pr = hooks.focused_probe({'sampling': samp, 'wavelength': 0.0197, 'dtype': numpy.float32},
                             {'conv_angle': 20.0, 'defocus': 100.0})
    assert abs(numpy.sum(numpy.abs(pr.data)**2) - 1) < 1e-5
    with pytest.raises(ValueError, match='conv_angle'):
        hooks.focused_probe({'sampling': samp, 'wavelength': 0.02, 'dtype': numpy.float32}, {'defocus': 1.0})


def test_random_object_hook_is_seeded():
    osamp = ObjectSampling((8, 9), (0.2, 0.2))
    a = hooks.random_object({'sampling': osamp, 'slices': [10., 10.], 'dtype': numpy.float32, 'seed': 5}, {})
    b = hooks.random_object({'sampling': osamp, 'slices': [10., 10.], 'dtype': numpy.float32, 'seed': 5}, {})
    assert a.data.shape == (2, 8, 9) and a.data.dtype == numpy.complex64
    numpy.testing.assert_array_equal(a.data, b.data)
    assert numpy.allclose(numpy.abs(a.data), 1) and 0 < numpy.abs(numpy.angle(a.data)).max() < 1e-4


def test_empad_round_trip_and_loader(tmp_path):
    """EMPAD frames are 130 x 128 with two junk rows and a flipped reciprocal y
    (io/empad.py:115-212); the loader reads metadata JSON, ifftshifts and masks."""
    import json
    rng = numpy.random.default_rng(0)
    arr = rng.random((3, 4, 128, 128)).astype(numpy.float32)
    raw = tmp_path / 'scan_x4_y3.raw'
    hooks.save_4d_empad(arr, raw)
    assert raw.stat().st_size == 3 * 4 * 130 * 128 * 4
    numpy.testing.assert_array_equal(hooks.load_4d_empad(raw), arr)
    meta = {'file_type': 'pyMultislicer_metadata', 'name': 'sim', 'raw_filename': raw.name, 'voltage': 300e3,
            'conv_angle': 25.0, 'defocus': 1.5e-9, 'diff_step': 1.0, 'scan_rotation': 0.0, 'scan_shape': [4, 3],
            'scan_fov': [1e-9, 1e-9], 'scan_step': [0.6e-10, 0.5e-10]}
    mj = tmp_path / 'meta.json'
    mj.write_text(json.dumps(meta))
    d = hooks.load_empad(None, {'path': str(mj)})
    numpy.testing.assert_array_equal(d['patterns'], numpy.fft.ifftshift(arr, axes=(-1, -2)))
    assert d['probe_hook'] == {'type': 'focused', 'conv_angle': 25.0, 'defocus': 15.0}
    assert d['scan_hook']['shape'] == (3, 4) and numpy.allclose(d['scan_hook']['step_size'], (0.5, 0.6))
    assert d['mask'][0, 0] == 1 and d['mask'][64, 64] == 0   # corner-origin: the border rows are masked
    assert abs(d['sampling'].extent[0] - d['wavelength'] / 1e-3) < 1e-9


def test_preprocessing_hooks():
    rng = numpy.random.default_rng(1)
    pats = rng.random((4, 5, 8, 8)).astype(numpy.float32)
    raw = {'patterns': pats.copy(), 'mask': numpy.ones((8, 8), numpy.float32), 'seed': 3,
           'scan_hook': {'type': 'raster', 'shape': (4, 5), 'step_size': 1.0}}
    out = hooks.crop_data(dict(raw), {'crop': (1, 3, None, 4)})
    assert out['patterns'].shape == (2, 4, 8, 8) and out['scan_hook']['shape'] == (2, 4)
    a = hooks.add_poisson_noise(dict(raw, patterns=pats.copy()), {'scale': 1e3})['patterns']
    b = hooks.add_poisson_noise(dict(raw, patterns=pats.copy()), {'scale': 1e3})['patterns']
    numpy.testing.assert_array_equal(a, b)                        # seeded 'poisson_noise' stream
    assert a.dtype == numpy.float32 and abs(a.mean() / (pats.mean() * 1e3) - 1) < 0.02
    assert numpy.abs(a - numpy.round(a)).max() < 0.01             # counts + N(0, 1e-3)
    # drop_nans: a pattern that is all NaN goes, with its scan position
    data = Patterns(pats.copy(), numpy.ones((8, 8), numpy.float32))
    data.patterns[1, 2] = numpy.nan
    st = NS(scan=numpy.arange(40, dtype=numpy.float32).reshape(4, 5, 2), tilt=None)
    d2, s2 = hooks.drop_nan_patterns({'data': data, 'state': st}, {})
    assert d2.patterns.shape == (19, 8, 8) and s2.scan.shape == (19, 2)
    assert not numpy.any(numpy.all(s2.scan == [14., 15.], axis=-1))


def test_diffraction_align_centres_the_mean_pattern():
    """A disc displaced by (+2, -1) px from zero frequency is moved back: the mean
    pattern's centre of mass ends within a small fraction of a pixel of zero."""
    n = 32
    f = numpy.fft.fftfreq(n, 1.0 / n)
    ky, kx = numpy.meshgrid(f, f, indexing='ij')
    disc = ((ky - 2)**2 + (kx + 1)**2 <= 16).astype(numpy.float32)
    pats = numpy.broadcast_to(disc, (3, 3, n, n)).copy()
    data = Patterns(pats, numpy.ones((n, n), numpy.float32))
    d, _ = hooks.diffraction_align({'data': data, 'state': NS(), 'dtype': numpy.float32}, {})
    m = d.patterns.mean(axis=(0, 1))
    com = [numpy.sum(k * m) / numpy.sum(m) for k in (ky, kx)]
    assert abs(com[0]) < 0.05 and abs(com[1]) < 0.05


def _state(S=2, M=1, n=16):
    rng = numpy.random.default_rng(2)
    psamp = Sampling((n, n), sampling=(0.2, 0.2))
    scan = hooks.make_raster_scan((4, 4), 0.6).astype(numpy.float32)
    osamp = ObjectSampling.from_scan(scan, psamp.sampling, psamp.extent / 2. + 5 * psamp.sampling)
    obj = numpy.exp(0.1j * rng.normal(size=(S, *osamp.shape))).astype(numpy.complex64)
    probe = (rng.normal(size=(M, n, n)) + 1j * rng.normal(size=(M, n, n))).astype(numpy.complex64)
    return ReconsState(iter=IterState(1, 3, 3), wavelength=0.0197, probe=ProbeState(psamp, probe),
                       object=ObjectState(osamp, obj, numpy.full(S, 20.0, numpy.float32)), scan=scan,
                       progress=ProgressState(iters=numpy.array([1, 2, 3], numpy.uint64),
                                              detector_errors=numpy.array([3., 2., 1.])))


def test_prepare_for_engine_modes_padding_and_slices():
    from phaser_amd.execute import prepare_for_engine
    st = _state(S=4)
    st.object.thicknesses = numpy.full(4, 10.0, numpy.float32)
    pats = Patterns(numpy.ones((4, 4, 16, 16), numpy.float32), numpy.ones((16, 16), numpy.float32))
    st.scan = st.scan * 1.5          # the scan outgrows the object: padded with vacuum (1)
    eng = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', probe_modes=3, slices=[20., 20.])
    _p, out = prepare_for_engine(pats, st, eng)
    assert out.probe.data.shape == (3, 16, 16)
    # unit-norm Hermite modes scaled by sqrt(power): 0.7 for the base, 1 - 0.7 for each
    # further one (make_hermetian_modes, optics.py:54-94)
    powers = numpy.sum(numpy.abs(out.probe.data)**2, axis=(1, 2))
    numpy.testing.assert_allclose(powers, [0.7, 0.3, 0.3], rtol=1e-4)
    assert out.object.data.shape[0] == 2 and numpy.allclose(out.object.thicknesses, 20.)
    assert all(out.object.sampling.shape > st.object.sampling.shape)
    out.object.sampling.check_scan(out.scan)
    # 4 x 10 A -> 2 x 20 A: each new slice centre lies midway between two old ones, so the
    # interpolated potential is exact and new slice k = old slices 2k and 2k+1 multiplied
    crop = tuple(slice(int(a), int(a) + s) for a, s in zip(numpy.round((st.object.sampling.corner -
                 out.object.sampling.corner) / 0.2), st.object.sampling.shape))
    o = st.object.data
    for k in range(2):
        numpy.testing.assert_allclose(out.object.data[k][crop], o[2 * k] * o[2 * k + 1], atol=2e-6)
    pad = numpy.ones(out.object.data.shape[1:], bool)
    pad[crop] = False
    assert numpy.allclose(out.object.data[:, pad], 1)          # vacuum outside the old grid


def test_reslicing_closed_forms():
    from phaser_amd.optics import resample_slices
    rng = numpy.random.default_rng(7)
    o = numpy.exp(0.2j * rng.normal(size=(3, 5, 6))).astype(numpy.complex64)
    numpy.testing.assert_allclose(resample_slices(o, [10., 10., 10.], [])[0], numpy.prod(o, axis=0), rtol=1e-5)
    one = o[:1]
    split = resample_slices(one, [], [10., 30.])
    numpy.testing.assert_allclose(split[0], numpy.exp(numpy.log(one[0]) * 0.25), rtol=1e-5)
    numpy.testing.assert_allclose(numpy.prod(split, axis=0), one[0], rtol=1e-5)


def test_sim_shape_resampling():
    from phaser_amd.execute import prepare_for_engine
    st = _state(n=16)
    pats = Patterns(numpy.ones((4, 4, 16, 16), numpy.float32), numpy.ones((16, 16), numpy.float32))
    eng = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', sim_shape=(32, 32))
    p2, out = prepare_for_engine(pats, st, eng)
    assert out.probe.data.shape == (1, 32, 32) and p2.patterns.shape == (4, 4, 32, 32)
    assert numpy.allclose(out.probe.sampling.extent, st.probe.sampling.extent)   # pad_crop keeps the extent


def test_checkpoint_round_trip(tmp_path):
    from phaser_amd.io import read_state, write_state
    st = _state(S=3, M=2)
    st.object.sampling.region_min = numpy.array([-1., -1.])
    st.object.sampling.region_max = numpy.array([1., 1.])
    path = write_state(st, tmp_path / 'iter003.npz')
    back = read_state(path)
    numpy.testing.assert_array_equal(back.object.data, st.object.data)
    numpy.testing.assert_array_equal(back.probe.data, st.probe.data)
    numpy.testing.assert_array_equal(back.scan, st.scan.astype(numpy.float64))
    numpy.testing.assert_allclose(back.object.sampling.corner, st.object.sampling.corner)
    numpy.testing.assert_allclose(back.object.sampling.region_max, [1., 1.])
    assert back.iter.total_iter == 3 and list(back.progress.detector_errors) == [3., 2., 1.]
    assert back.wavelength == pytest.approx(0.0197)
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match='npz'):
            write_state(st, tmp_path / 'x.h5')


def test_save_and_patience_observers(tmp_path):
    from phaser_amd.execute import PatienceObserver, SaveObserver
    from phaser_amd.plan import Flag, SaveOptions
    from phaser_amd.state import EarlyTermination
    st = _state()
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=4, save=Flag(every=2),
                                  save_options=SaveOptions(out_dir=str(tmp_path / '{name}')))
    so = SaveObserver()
    so.init_engine(st, recons_name='r', plan=plan)
    for i in range(1, 5):
        st.iter.total_iter = i
        st.iter.engine_iter = i
        so.update_iteration(st, i, 4)
    so.finish_engine(st)
    so.close()
    files = sorted(p.name for p in (tmp_path / 'r').iterdir())
    assert 'finished' in files and 'iter002.npz' in files and 'iter004.npz' in files and 'iter001.npz' not in files
    po = PatienceObserver(2)
    st.progress.detector_errors = numpy.array([3.])
    po.update_iteration(st, 1, 9)
    for e in (4., 5.):
        st.progress.detector_errors = numpy.append(st.progress.detector_errors, e)
        if e == 5.:
            with pytest.raises(EarlyTermination):
                po.update_iteration(st, 3, 9)
        else:
            po.update_iteration(st, 2, 9)
