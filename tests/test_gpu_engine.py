"""GPU parity of the conventional engines against the CPU oracle on the same seeded
inputs.  Metric: relative L2 error ||a - b|| / ||b|| of the object, the probe and the
per-iteration detector errors.

Every case also runs the oracle's FP64 twin and states the error of both FP32 results
against FP64.  Two bars:
  * well-conditioned runs (the FP32 oracle agrees with its FP64 twin to < 1e-5): the GPU
    must agree with the FP32 oracle within 1e-4 (the parity bar);
  * ill-conditioned runs (FP32 rounding is amplified by near-null directions of the
    object/probe problem, or the run diverges): the GPU's error against FP64 must be no
    more than 3x the FP32 oracle's own error against FP64 (+1e-5).
"""
import numpy
import pytest

import oracle.engine as oeng
import oracle.noise_models as onoise
from tests.helpers import rel_err, to_oracle_state, oracle_simulate

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _problem(**kw):
    from phaser_amd.synthetic import SyntheticConfig, make_dataset
    cfg = SyntheticConfig(**kw)
    state, pats, truth = make_dataset(cfg, simulate=oracle_simulate)
    return cfg, state, pats


def _run_oracle(state, pats, dtype, **kw):
    o = to_oracle_state(state, dtype=dtype)
    oeng.run_engine(o, pats.patterns.astype(dtype), pats.pattern_mask.astype(dtype), dtype=dtype, **kw)
    return o


def _run_all(state, pats, *, solver, niter, grouping, use_graphs=False, position=None, beta=1.0, noise='amplitude'):
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, MomentumPositionSolverPlan, SOLVERS
    splan = SOLVERS[solver](beta_object=beta, beta_probe=beta)
    plan = ConventionalEnginePlan(noise_model=noise, solver=splan, niter=niter, grouping=grouping,
                                  position_solver=MomentumPositionSolverPlan(**position) if position else None,
                                  update_positions=bool(position))
    g = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                    'recons_name': 't'}, plan, use_graphs=use_graphs).to_numpy()
    nm = onoise.AnscombeNoiseModel() if noise == 'anscombe' else onoise.AmplitudeNoiseModel()
    kw = dict(solver=solver, noise_model=nm, niter=niter, grouping=grouping, seed=7, beta_object=beta,
              beta_probe=beta, update_positions=bool(position))
    out = []
    for dt in (numpy.float32, numpy.float64):
        pos = oeng.MomentumPositionSolver(**position) if position else None
        out.append(_run_oracle(state, pats, dt, position_solver=pos, **kw))
    return g, out[0], out[1]


def _check(g, o32, o64, strict=True):
    res = {}
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('errors', lambda s: s.progress.detector_errors)):
        res[name] = (rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64)))
        print(f"{name}: gpu-vs-fp32 {res[name][0]:.2e}  gpu-vs-fp64 {res[name][1]:.2e}  "
              f"fp32-vs-fp64 {res[name][2]:.2e}", flush=True)
    assert len(g.progress.detector_errors) == len(o32.progress.detector_errors)
    for name, (go, g64, o64_) in res.items():
        if strict:
            assert o64_ < 1e-5, f"{name}: case is not well-conditioned ({o64_:.2e})"
            assert go < TOL, name
        else:
            assert g64 <= 3 * o64_ + 1e-5, name
    return res


@pytest.mark.parametrize('n', [64, 128, 192, 256])
def test_gpu_forward_simulation_matches_oracle(n):
    """The synthetic patterns the benchmarks generate on the GPU (phx_shift_probes,
    phx_forward_slice and phx_exit_intensity: the per-slice forward model and the mode-sum
    of the exit intensity, written in natural order) against the oracle's FP64 forward
    model, at every plane side the engines take, with modes and three slices."""
    from phaser_amd.synthetic import (SyntheticConfig, make_ground_truth, simulate_patterns_gpu,
                                      thicknesses)
    cfg = SyntheticConfig(n=n, n_modes=2, n_slices=3, scan_shape=(2, 3), scan_step_px=9.0, jitter_px=0.5,
                          seed=n)
    truth, osamp, scan, probe, wavelength, psamp = make_ground_truth(cfg)
    thick = thicknesses(cfg)
    g = simulate_patterns_gpu(truth, osamp, scan, probe, wavelength, psamp, thick)
    o = oracle_simulate(truth, osamp, scan, probe, wavelength, psamp, thick)
    e = rel_err(g, o)
    print(f"n={n}: simulated patterns vs FP64 oracle {e:.2e}", flush=True)
    assert g.shape == o.shape and e < 1e-5


def test_lsqml_one_iteration_multislice_small():
    _, state, pats = _problem(n=32, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5,
                              seed=1, defocus=30.0)
    _check(*_run_all(state, pats, solver='lsqml', niter=1, grouping=12))


def test_epie_one_iteration_multislice_small():
    _, state, pats = _problem(n=32, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5,
                              seed=2, defocus=30.0)
    _check(*_run_all(state, pats, solver='epie', niter=1, grouping=12))


def test_epie_config1_ten_iterations():
    """Config 1 (single-slice ePIE, 64^2 probe, 16x16 scan step 4 px with +-0.5 px jitter,
    Poisson 1e5, 10 iterations) with groups of 16.  With groups of 64 and beta = 1 this
    ePIE diverges on this data (see the next test), so the parity bar is applied here."""
    _, state, pats = _problem(n=64, n_modes=1, n_slices=1, scan_shape=(16, 16), scan_step_px=4.0, jitter_px=0.5,
                              seed=0)
    _check(*_run_all(state, pats, solver='epie', niter=10, grouping=16))


def test_epie_config1_default_groups_diverging():
    """Same data, reference defaults (groups of 64, beta 1): the detector error grows by
    ~1e5 in 10 iterations, so only the FP64-relative bar is meaningful."""
    _, state, pats = _problem(n=64, n_modes=1, n_slices=1, scan_shape=(16, 16), scan_step_px=4.0, jitter_px=0.5,
                              seed=0)
    _check(*_run_all(state, pats, solver='epie', niter=10, grouping=64), strict=False)


def test_epie_config1_beta02_groups64():
    _, state, pats = _problem(n=64, n_modes=1, n_slices=1, scan_shape=(16, 16), scan_step_px=4.0, jitter_px=0.5,
                              seed=0)
    _check(*_run_all(state, pats, solver='epie', niter=10, grouping=64, beta=0.2))


def test_lsqml_single_slice_config1():
    """Single-slice LSQ-ML on config 1's data: FP32 rounding drifts along the object/probe
    ambiguity outside the scanned region (~1e-3 vs FP64), so the FP64-relative bar."""
    _, state, pats = _problem(n=64, n_modes=1, n_slices=1, scan_shape=(16, 16), scan_step_px=4.0, jitter_px=0.5,
                              seed=0)
    _check(*_run_all(state, pats, solver='lsqml', niter=10, grouping=64), strict=False)


def test_lsqml_multislice_modes_with_graphs():
    """Reduced 8-slice LSQ-ML configuration (4 modes, 4 slices, jitter-free raster: start
    indices hit exact .5 ties), whole iterations captured in a HIP graph and replayed."""
    _, state, pats = _problem(n=64, n_modes=4, n_slices=4, scan_shape=(12, 12), scan_step_px=6.0, jitter_px=0.0,
                              seed=4)
    _check(*_run_all(state, pats, solver='lsqml', niter=5, grouping=48, use_graphs=True))


def test_lsqml_graphs_probe_update_switching_on():
    """Graphs captured both without the probe update (iterations 0-1: the group steps run
    on the caller's stream alone) and with it (iterations 2-3: probe-side kernels on the
    second stream), replayed in turn, against the oracle with the same flag."""
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    _, state, pats = _problem(n=64, n_modes=2, n_slices=3, scan_shape=(8, 8), scan_step_px=6.0, jitter_px=0.5,
                              seed=5)
    flag = {'after': 1}
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5),
                                  niter=4, grouping=16, update_probe=flag)
    g = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                    'recons_name': 't'}, plan, use_graphs=True).to_numpy()
    kw = dict(solver='lsqml', noise_model=onoise.AmplitudeNoiseModel(), niter=4, grouping=16, seed=7,
              beta_object=0.5, beta_probe=0.5, update_probe=flag)
    o32 = _run_oracle(state, pats, numpy.float32, **kw)
    o64 = _run_oracle(state, pats, numpy.float64, **kw)
    _check(g, o32, o64)


@pytest.mark.parametrize('solver', ['lsqml', 'epie'])
def test_n128_multislice_modes(solver):
    """128^2 planes (the benchmark's size, whose transforms use 8 lanes per row): 3 slices,
    2 modes, 2 iterations of the fused LSQ-ML sweeps / ePIE, against the oracle."""
    _, state, pats = _problem(n=128, n_modes=2, n_slices=3, scan_shape=(4, 4), scan_step_px=6.0, jitter_px=0.5,
                              seed=8)
    _check(*_run_all(state, pats, solver=solver, niter=2, grouping=8, beta=0.5))


def test_band_limit_column_skip_matches(monkeypatch):
    """At 128^2 the band-limited propagators let waves 6-9 of every plane skip their
    column passes (the product is zero there). The fused LSQ-ML sweeps give the same
    result with and without the skip: the skipped product is exactly zero.  Every FP32
    sum of the LSQ-ML step is ordered (the modes' intensities in mode order, the
    object-magnitude sum and the probe spectra in position order, the update gathered
    per tile), so with four modes two runs agree to the last bit and the skip must not
    change one either (the FP64 step-size sums, added in any order, are read as FP32
    scalars)."""
    import phaser_amd.device as dev
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    _, state, pats = _problem(n=128, n_modes=4, n_slices=3, scan_shape=(4, 4), scan_step_px=6.0, jitter_px=0.5,
                              seed=8)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5),
                                  niter=2, grouping=8)
    args = {'data': pats, 'dtype': numpy.float32, 'seed': 7, 'observer': None, 'recons_name': 't'}
    masks = []
    real = dev.zero_column_waves

    def spy(props, n):
        masks.append(real(props, n))
        return masks[-1]
    monkeypatch.setattr(dev, 'zero_column_waves', spy)
    a = run_engine(dict(args, state=state.copy()), plan, use_graphs=False).to_numpy()
    a2 = run_engine(dict(args, state=state.copy()), plan, use_graphs=False).to_numpy()
    monkeypatch.setattr(dev, 'zero_column_waves', lambda props, n: 0)
    b = run_engine(dict(args, state=state.copy()), plan, use_graphs=False).to_numpy()
    assert masks and masks[0] == 0b1111000000

    def rel(x, y):
        return float(numpy.abs(x - y).max() / numpy.abs(y).max())
    noise = max(rel(a.object.data, a2.object.data), rel(a.probe.data, a2.probe.data))
    e = max(rel(a.object.data, b.object.data), rel(a.probe.data, b.probe.data))
    print(f"run-to-run {noise:.2e}, skip vs no skip {e:.2e}", flush=True)
    assert noise == 0.0
    assert e == 0.0


def test_lsqml_deep_stack_block_sums():
    """34 slices: more than the fused backward sweep keeps per-wave alpha partials for in
    LDS (32), so every slice step sums its alpha terms across the block instead; 2
    iterations against the oracle.  A stack this deep is ill-conditioned in FP32 (the
    FP32 oracle's probe differs from FP64 by ~2e-3), so the FP64-relative bar."""
    _, state, pats = _problem(n=32, n_modes=1, n_slices=34, scan_shape=(6, 6), scan_step_px=4.0, jitter_px=0.5,
                              seed=21)
    _check(*_run_all(state, pats, solver='lsqml', niter=2, grouping=36, beta=0.5), strict=False)


def test_probe_step_strips_match_single_launch(monkeypatch):
    """The probe step in three strip launches (beside the object update) against the
    one-workgroup-per-mode launch after the reduction, at 128^2 and 64^2 with graphs.
    The path is deterministic (ordered sums), so with four modes two runs must agree to
    the last bit; the two probe-step forms run the same transforms in different launch
    shapes and must agree to FP32 rounding."""
    import phaser_amd.device as dev
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    for n, scan, grouping in ((128, (4, 4), 8), (64, (8, 8), 16)):
        _, state, pats = _problem(n=n, n_modes=4, n_slices=3, scan_shape=scan, scan_step_px=6.0, jitter_px=0.5,
                                  seed=8)
        plan = ConventionalEnginePlan(noise_model='amplitude',
                                      solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5), niter=3,
                                      grouping=grouping)
        args = {'data': pats, 'dtype': numpy.float32, 'seed': 7, 'observer': None, 'recons_name': 't'}
        monkeypatch.setattr(dev, 'PROBE_STRIPS_MAX_N', 128)
        a = run_engine(dict(args, state=state.copy()), plan, use_graphs=True).to_numpy()
        a2 = run_engine(dict(args, state=state.copy()), plan, use_graphs=True).to_numpy()
        monkeypatch.setattr(dev, 'PROBE_STRIPS_MAX_N', 0)
        b = run_engine(dict(args, state=state.copy()), plan, use_graphs=True).to_numpy()

        def rel(x, y):
            return float(numpy.abs(x - y).max() / numpy.abs(y).max())
        noise = max(rel(a.object.data, a2.object.data), rel(a.probe.data, a2.probe.data))
        e = max(rel(a.object.data, b.object.data), rel(a.probe.data, b.probe.data))
        print(n, f'run-to-run {noise:.2e}, strips vs single {e:.2e}', flush=True)
        assert noise == 0.0
        assert e <= 1e-6


def test_device_resident_patterns():
    """Patterns handed over as a float32 tensor already on the GPU (no host copy) give the
    same reconstruction as host patterns, for the conventional and the gradient engine."""
    import torch
    from dataclasses import replace
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.engines import gradient as geng
    from phaser_amd.plan import ConventionalEnginePlan, GradientEnginePlan, SOLVERS
    _, state, pats = _problem(n=64, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=6.0, jitter_px=0.5,
                              seed=9)
    dpats = replace(pats, patterns=torch.from_numpy(numpy.ascontiguousarray(pats.patterns)).cuda())

    def rel(x, y):
        return float(numpy.abs(x - y).max() / numpy.abs(y).max())
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5),
                                  niter=2, grouping=12)
    gplan = GradientEnginePlan(noise_model='amplitude', niter=2, grouping=12,
                               solvers={'object': {'type': 'sgd', 'learning_rate': 1.0, 'momentum': 0.9}})
    for run, pl in ((run_engine, plan), (geng.run_engine, gplan)):
        out = []
        for p in (pats, dpats):
            out.append(run({'data': p, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                            'recons_name': 't'}, pl))
        a, b = (o.to_numpy() if hasattr(o, 'to_numpy') else o for o in out)
        oa = a.object.data if not hasattr(a.object.data, 'cpu') else a.object.data.cpu().numpy()
        ob = b.object.data if not hasattr(b.object.data, 'cpu') else b.object.data.cpu().numpy()
        assert rel(oa, ob) < 1e-5


def test_n192_multislice_lsqml():
    """192^2 planes (config 5's probe size: DFT-48 per lane, transposes through a global
    scratch plane): 2 slices, 2 modes, 2 iterations of the fused LSQ-ML sweeps, against
    the oracle."""
    _, state, pats = _problem(n=192, n_modes=2, n_slices=2, scan_shape=(3, 3), scan_step_px=8.0, jitter_px=0.5,
                              seed=12)
    _check(*_run_all(state, pats, solver='lsqml', niter=2, grouping=9, beta=0.5))


@pytest.mark.parametrize('slices,modes,graphs', [(3, 2, False), (3, 2, True), (1, 1, False)])
def test_n256_lsqml_split_passes(slices, modes, graphs):
    """256^2 planes (a plane is too large for one workgroup, so every 2-D transform is a
    row launch and a column launch through HBM, spectra in natural order): LSQ-ML with
    object and probe updates, two iterations against the oracle, with and without graphs,
    and the single-slice case.  The single-slice, single-mode probe is ill-conditioned in
    FP32 (the FP32 oracle's probe is ~1e-4 from FP64), so that case takes the
    FP64-relative bar."""
    _, state, pats = _problem(n=256, n_modes=modes, n_slices=slices, scan_shape=(3, 3), scan_step_px=10.0,
                              jitter_px=0.5, seed=14)
    _check(*_run_all(state, pats, solver='lsqml', niter=2, grouping=9, beta=0.5, use_graphs=graphs),
           strict=slices > 1)


def test_n256_folded_passes_equal_per_slice_launches():
    """At 256^2 the group forward and backward sweeps fold each slice step's finishing
    row pass into the next step's first: the same arithmetic as the per-slice launches
    (the finished wave is stored and used from registers instead of re-read), so LSQ-ML
    with probe updates gives the same bits either way."""
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    _, state, pats = _problem(n=256, n_modes=2, n_slices=3, scan_shape=(3, 3), scan_step_px=10.0,
                              jitter_px=0.5, seed=17)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5),
                                  niter=2, grouping=5)
    args = {'data': pats, 'dtype': numpy.float32, 'seed': 7, 'observer': None, 'recons_name': 't'}
    a = run_engine(dict(args, state=state.copy()), plan, use_graphs=False).to_numpy()
    b = run_engine(dict(args, state=state.copy()), plan, use_graphs=False, fused=False).to_numpy()
    assert numpy.isfinite(a.object.data).all()
    assert numpy.array_equal(a.object.data, b.object.data)
    assert numpy.array_equal(a.probe.data, b.probe.data)
    numpy.testing.assert_array_equal(a.progress.detector_errors, b.progress.detector_errors)


def test_n256_epie_split_passes():
    """ePIE at 256^2 (split row and column launches, object update gathered per tile):
    three slices, two modes, two iterations against the oracle."""
    _, state, pats = _problem(n=256, n_modes=2, n_slices=3, scan_shape=(3, 3), scan_step_px=10.0,
                              jitter_px=0.5, seed=15)
    _check(*_run_all(state, pats, solver='epie', niter=2, grouping=9, beta=0.5))


def test_n256_lsqml_position_correction():
    """LSQ-ML with momentum position correction at 256^2 (the position step runs in split
    passes through the farfield slots, before slice 0's probe spectra go there): scan,
    object and probe against the oracle over two iterations."""
    _, state, pats = _problem(n=256, n_modes=2, n_slices=2, scan_shape=(3, 3), scan_step_px=10.0,
                              jitter_px=0.5, seed=16)
    pos = {'step_size': 1e-2, 'max_step_size': 0.1, 'momentum': 0.9}
    g, o32, o64 = _run_all(state, pats, solver='lsqml', niter=2, grouping=9, beta=0.5, position=pos)
    _check(g, o32, o64)
    # the scan update on the FP64-relative bar, as at the smaller sizes
    d_g, d_o, d_64 = g.scan - state.scan, o32.scan - state.scan, o64.scan - state.scan
    e_g64, e_o64 = rel_err(d_g, d_64), rel_err(d_o, d_64)
    print(f"scan update: gpu-vs-fp32 {rel_err(d_g, d_o):.2e} gpu-vs-fp64 {e_g64:.2e} fp32-vs-fp64 {e_o64:.2e}",
          flush=True)
    assert numpy.abs(d_o).max() > 0
    assert e_g64 <= 3 * e_o64 + 1e-5


def test_n256_gradient_recompute_mode_raises():
    """At 256^2 the gradient engine keeps every slice's wave; its recompute mode has no
    split-pass kernels and is refused before anything runs."""
    from phaser_amd.engines import gradient as geng
    from phaser_amd.plan import GradientEnginePlan
    _, state, pats = _problem(n=256, n_modes=1, n_slices=1, scan_shape=(2, 2), scan_step_px=10.0, jitter_px=0.0,
                              seed=3)
    args = {'data': pats, 'dtype': numpy.float32, 'seed': 7, 'observer': None, 'recons_name': 't'}
    gplan = GradientEnginePlan(noise_model='amplitude', niter=1, grouping=4,
                               solvers={'object': {'type': 'sgd', 'learning_rate': 1.0}})
    with pytest.raises(NotImplementedError):
        geng.run_engine(dict(args, state=state.copy()), gplan, recompute=True)


def test_lsqml_anscombe_multislice():
    _, state, pats = _problem(n=32, n_modes=2, n_slices=2, scan_shape=(8, 8), scan_step_px=3.0, jitter_px=0.5,
                              seed=6, defocus=30.0)
    _check(*_run_all(state, pats, solver='lsqml', niter=3, grouping=16, noise='anscombe'))


def test_lsqml_position_correction():
    _, state, pats = _problem(n=32, n_modes=2, n_slices=2, scan_shape=(8, 8), scan_step_px=3.0, jitter_px=0.5,
                              seed=5, defocus=30.0)
    pos = {'step_size': 1e-2, 'max_step_size': 0.1, 'momentum': 0.9}
    g, o32, o64 = _run_all(state, pats, solver='lsqml', niter=3, grouping=16, position=pos)
    _check(g, o32, o64)
    # the position step (a ratio of sums with heavy cancellation) drifts ~2e-3 between
    # FP32 and FP64 even in the oracle, so it gets the FP64-relative bar
    d_g, d_o, d_64 = g.scan - state.scan, o32.scan - state.scan, o64.scan - state.scan
    e_g64, e_o64 = rel_err(d_g, d_64), rel_err(d_o, d_64)
    print(f"scan update: gpu-vs-fp32 {rel_err(d_g, d_o):.2e} gpu-vs-fp64 {e_g64:.2e} fp32-vs-fp64 {e_o64:.2e}",
          flush=True)
    assert numpy.abs(d_o).max() > 0
    assert e_g64 <= 3 * e_o64 + 1e-5


def test_epie_position_correction():
    """ePIE with momentum position correction against the oracle: ePIE's group step always
    computes the position gradients (solvers.py:489-502; its run_iteration never forwards
    update_position), and the engine loop applies them (run.py:95-110)."""
    _, state, pats = _problem(n=32, n_modes=2, n_slices=2, scan_shape=(8, 8), scan_step_px=3.0, jitter_px=0.5,
                              seed=5, defocus=30.0)
    pos = {'step_size': 1e-2, 'max_step_size': 0.1, 'momentum': 0.9}
    g, o32, o64 = _run_all(state, pats, solver='epie', niter=3, grouping=16, position=pos, beta=0.5)
    _check(g, o32, o64)
    d_g, d_o, d_64 = g.scan - state.scan, o32.scan - state.scan, o64.scan - state.scan
    e_g64, e_o64 = rel_err(d_g, d_64), rel_err(d_o, d_64)
    print(f"scan update: gpu-vs-fp32 {rel_err(d_g, d_o):.2e} gpu-vs-fp64 {e_g64:.2e} fp32-vs-fp64 {e_o64:.2e}",
          flush=True)
    assert numpy.abs(d_o).max() > 0
    assert e_g64 <= 3 * e_o64 + 1e-5


@pytest.mark.parametrize('solver,slices', [('lsqml', 3), ('epie', 3), ('lsqml', 5)])
def test_fused_sweeps_and_deferred_updates_equal_per_slice_launches(solver, slices):
    """Whole-sweep kernels (one launch for the forward model, one for the LSQ-ML backward
    sweep, every object update after the sweep) give the per-slice launch sequence's
    result: the sweep reads only pre-update cutouts (solvers.py:246-260)."""
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, MomentumPositionSolverPlan
    _, state, pats = _problem(n=32, n_modes=2, n_slices=slices, scan_shape=(6, 6), scan_step_px=3.0,
                              jitter_px=0.5, seed=5)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=solver, niter=2, grouping=12,
                                  position_solver=MomentumPositionSolverPlan(step_size=1e-2),
                                  update_positions=True)
    variants = [dict(keep_chi=False, fused=False), dict(keep_chi=True, fused=False), dict(keep_chi=True, fused=True)]
    out = []
    for kw in variants:
        args = {'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 3, 'observer': None,
                'recons_name': 't'}
        out.append(run_engine(args, plan, use_graphs=False, **kw).to_numpy())
    ref = out[0]
    for kw, o in zip(variants[1:], out[1:]):
        eo, ep, es = (rel_err(o.object.data, ref.object.data), rel_err(o.probe.data, ref.probe.data),
                      rel_err(o.scan, ref.scan))
        print(f"{kw}: object {eo:.2e} probe {ep:.2e} scan {es:.2e}", flush=True)
        assert eo < 1e-6 and ep < 1e-6 and es < 1e-6
        # the fused forward launch rounds differently from the per-slice launches (object
        # ~3e-7 apart after two ePIE iterations), which moves ePIE's second-iteration
        # error at the 1e-6 level
        numpy.testing.assert_allclose(o.progress.detector_errors, ref.progress.detector_errors,
                                      rtol=1e-5 if solver == 'epie' else 1e-6)


def test_update_launch_clearing_next_group_matches_separate_clear(monkeypatch):
    """The object update that also zeroes the next group's accumulators
    (phx_lsqml_object_update_slices_begin) against phx_group_begin launched on its own
    before each group, with graphs (probe update switching on after the first
    iteration) and without.  The fold only moves a clear into another launch, and the
    path is deterministic (ordered sums), so with four modes both must agree to the last
    bit."""
    import phaser_amd.device as dev
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    _, state, pats = _problem(n=128, n_modes=4, n_slices=3, scan_shape=(4, 4), scan_step_px=6.0, jitter_px=0.5,
                              seed=8)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5),
                                  niter=3, grouping=4, update_probe={'after': 1})
    args = {'data': pats, 'dtype': numpy.float32, 'seed': 7, 'observer': None, 'recons_name': 't'}

    def rel(x, y):
        return float(numpy.abs(x - y).max() / numpy.abs(y).max())
    for graphs in (True, False):
        monkeypatch.setattr(dev, 'FOLD_GROUP_BEGIN', True)
        a = run_engine(dict(args, state=state.copy()), plan, use_graphs=graphs).to_numpy()
        a2 = run_engine(dict(args, state=state.copy()), plan, use_graphs=graphs).to_numpy()
        monkeypatch.setattr(dev, 'FOLD_GROUP_BEGIN', False)
        b = run_engine(dict(args, state=state.copy()), plan, use_graphs=graphs).to_numpy()
        noise = max(rel(a.object.data, a2.object.data), rel(a.probe.data, a2.probe.data))
        e = max(rel(a.object.data, b.object.data), rel(a.probe.data, b.probe.data))
        print('graphs', graphs, f'run-to-run {noise:.2e}, fold vs separate {e:.2e}', flush=True)
        assert noise == 0.0
        assert e == 0.0
        numpy.testing.assert_allclose(a.progress.detector_errors, b.progress.detector_errors, rtol=1e-6)


def test_epie_runs_agree_to_the_last_bit():
    """ePIE's object update is gathered per object tile in group order (no float atomics),
    the modes' intensities are added in mode order and the probe spectra in position
    order, so two runs with four modes and three slices agree to the last bit."""
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan
    _, state, pats = _problem(n=64, n_modes=4, n_slices=3, scan_shape=(6, 6), scan_step_px=5.0, jitter_px=0.5,
                              seed=12)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='epie', niter=3, grouping=12)
    args = {'data': pats, 'dtype': numpy.float32, 'seed': 7, 'observer': None, 'recons_name': 't'}
    a = run_engine(dict(args, state=state.copy()), plan).to_numpy()
    b = run_engine(dict(args, state=state.copy()), plan).to_numpy()
    assert numpy.isfinite(a.object.data).all() and numpy.isfinite(a.probe.data).all()
    assert numpy.array_equal(a.object.data, b.object.data)
    assert numpy.array_equal(a.probe.data, b.probe.data)


@pytest.mark.parametrize('solver,per_group', [('lsqml', False), ('lsqml', True), ('epie', False)])
def test_nan_in_state_raises_reference_error(solver, per_group):
    """A NaN already in the object (in the last slice, which the LSQ-ML/ePIE dry runs skip,
    so the presolve stays finite): the reference's check after group 0 sees the whole
    object (solvers.py:109) and raises for group 0.  The GPU engine raises the same
    ValueError, with graphs and with an observer that takes every group's state."""
    from phaser_amd.engines.conventional import run_engine, Observer
    from phaser_amd.plan import ConventionalEnginePlan
    _, state, pats = _problem(n=32, n_modes=1, n_slices=2, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5,
                              seed=11)
    state.object.data = numpy.array(state.object.data)
    state.object.data[1, 2, 2] = numpy.nan
    o = to_oracle_state(state)
    with pytest.raises(ValueError) as oe:
        oeng.run_engine(o, pats.patterns, pats.pattern_mask, solver=solver, noise_model=onoise.AmplitudeNoiseModel(),
                        niter=2, grouping=12, seed=7)

    class EveryGroup(Observer):
        def __init__(self):
            self.groups = 0

        def update_group(self, state, force=False):
            self.groups += 1
    obs = EveryGroup() if per_group else None
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=solver, niter=2, grouping=12)
    with pytest.raises(ValueError) as ge:
        run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': obs,
                    'recons_name': 't'}, plan)
    print(f"oracle: {oe.value}; gpu: {ge.value}", flush=True)
    assert str(ge.value) == str(oe.value) == "NaN or inf encountered in object or probe, group 0"
    if per_group:
        assert obs.groups == 0     # raised before the observer saw group 0's state
