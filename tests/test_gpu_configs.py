"""The benchmark configurations' per-position shapes on the GPU against the oracle, on
scans small enough for the numpy oracle (seconds), plus full-size property checks.

  * config 2: 128^2 probe x 4 modes, 8 slices of 20 A, jitter-free raster at 6 px (exact
    .5 ties), LSQ-ML defaults, 10 iterations; the 1e-4 bar on the object over the scanned
    region (get_region_crop), the probe and the per-iteration errors, with the error of
    both FP32 results against the FP64 twin stated;
  * config 3: 128^2 x 6 modes, 16 slices, the gradient engine with position correction
    (a momentum SGD solver on the positions) against the oracle's gradient engine;
  * config 5: 192^2 x 8 modes, 24 slices, LSQ-ML on a 2 x 2 scan;
  * full size, configs 2 and 3: one iteration of the whole problem is finite and lowers
    the detector error, and group 0's step equals the oracle's on that group;
  * full size, config 5: groups 0 and 1 of the 24-slice problem step as the oracle's.
"""
import numpy
import pytest

import oracle.engine as oeng
import oracle.gradient as og
import oracle.noise_models as onoise
from tests.helpers import oracle_simulate, rel_err, to_oracle_state

pytestmark = pytest.mark.gpu


def _problem(**kw):
    from phaser_amd.synthetic import SyntheticConfig, make_dataset
    state, pats, truth = make_dataset(SyntheticConfig(**kw), simulate=oracle_simulate)
    return state, pats


def _crop(state_like, arr):
    """The scanned region of an object (ObjectSampling.get_region_crop, object.py:367-378)."""
    from phaser_amd.sampling import ObjectSampling
    s = state_like.object.sampling
    osamp = ObjectSampling(s.shape, s.sampling, s.corner, s.region_min, s.region_max)
    return numpy.asarray(arr)[(Ellipsis, *osamp.get_region_crop())]


def _report(tag, g, o32, o64, crop_state):
    res = {}
    for name, f in (('object (scanned region)', lambda s: _crop(crop_state, s.object.data)),
                    ('probe', lambda s: s.probe.data), ('errors', lambda s: s.progress.detector_errors)):
        a, b, c = rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64))
        print(f"{tag} {name}: gpu-vs-fp32 {a:.2e} gpu-vs-fp64 {b:.2e} fp32-vs-fp64 {c:.2e}", flush=True)
        res[name] = (a, b, c)
    return res


def test_config2_shape_lsqml_ten_iterations():
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan
    state, pats = _problem(n=128, n_modes=4, n_slices=8, scan_shape=(3, 3), scan_step_px=6.0, jitter_px=0.0, seed=2)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=10, grouping=9)
    g = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                    'recons_name': 't'}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        oeng.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), dtype=dt, solver='lsqml',
                        noise_model=onoise.AmplitudeNoiseModel(), niter=10, grouping=9, seed=7)
        res.append(o)
    r = _report('config 2 shape', g, *res, state)
    for name, (a, b, c) in r.items():
        assert a < 1e-4, name
        assert b <= 3 * c + 1e-5, name


# the solvers of the reference's example gradient plan (examples/si_grad.yaml:37-50)
def _example_solvers():
    from phaser_amd.plan import AdamSolverPlan, SGDSolverPlan
    return {'object': AdamSolverPlan(learning_rate=5e-3, nesterov=True),
            'probe': AdamSolverPlan(learning_rate=0.1, nesterov=True),
            'positions': SGDSolverPlan(learning_rate=0.5, momentum=0.9)}


def test_config3_shape_gradient_engine_with_positions():
    from phaser_amd.engines.gradient import run_engine
    from phaser_amd.plan import GradientEnginePlan
    state, pats = _problem(n=128, n_modes=6, n_slices=16, scan_shape=(2, 2), scan_step_px=6.0, jitter_px=1.0,
                           seed=3, dose=5e4)
    state.probe.data = (state.probe.data / numpy.float32(numpy.sqrt(5e4))).astype(numpy.complex64)
    plan = GradientEnginePlan(noise_model='amplitude', niter=3, grouping=4, solvers=_example_solvers(),
                              update_positions=True)
    g = run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32}, plan).to_numpy()
    res = []
    nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        og.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), noise=nz, dtype=dt, niter=3,
                      grouping=4, seed=3, update_positions=True,
                      solvers=[og.Solver('adam', ['object'], learning_rate=5e-3, nesterov=True),
                               og.Solver('adam', ['probe'], learning_rate=0.1, nesterov=True),
                               og.Solver('sgd', ['positions'], learning_rate=0.5, momentum=0.9, nesterov=True)])
        res.append(o)
    o32, o64 = res
    r = _report('config 3 shape', g, o32, o64, state)
    d_g, d_32, d_64 = g.scan - state.scan, o32.scan - state.scan, o64.scan - state.scan
    es = (rel_err(d_g, d_32), rel_err(d_g, d_64), rel_err(d_32, d_64))
    print(f"config 3 shape scan update (max {numpy.abs(d_64).max():.2e} A): gpu-vs-fp32 {es[0]:.2e} "
          f"gpu-vs-fp64 {es[1]:.2e} fp32-vs-fp64 {es[2]:.2e}", flush=True)
    assert numpy.abs(d_32).max() > 0
    for name, (a, b, c) in list(r.items()) + [('scan', es)]:
        assert a < 1e-4 or b <= 3 * c + 1e-5, name


def test_config5_shape_lsqml_192():
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    state, pats = _problem(n=192, n_modes=8, n_slices=24, scan_shape=(2, 2), scan_step_px=7.0, jitter_px=0.0,
                           seed=5, dose=2e4)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS['lsqml'](beta_object=0.5, beta_probe=0.5),
                                  niter=2, grouping=4)
    g = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                    'recons_name': 't'}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        oeng.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), dtype=dt, solver='lsqml',
                        noise_model=onoise.AmplitudeNoiseModel(), niter=2, grouping=4, seed=7, beta_object=0.5,
                        beta_probe=0.5)
        res.append(o)
    r = _report('config 5 shape', g, *res, state)
    for name, (a, b, c) in r.items():
        assert a < 1e-4 if c < 1e-5 else b <= 3 * c + 1e-5, name


# ------------------------------------------------------------------ full size
def _full(cfg):
    from phaser_amd.synthetic import make_dataset
    return make_dataset(cfg)


def test_config2_full_size_iteration_and_first_group():
    """All 4096 positions of the 8-slice 518^2 problem: two LSQ-ML iterations are finite and
    lower the detector error (every group's error computed); then, from the GPU presolve's
    probe, group 0 of iteration 1 changes the object and probe as the oracle's lsqml_run
    does on that group (relative error of the updates)."""
    import torch
    import oracle.solvers as osol
    from oracle.optics import make_propagators as omake_props
    from phaser_amd.device import GroupWorkspace
    from phaser_amd.engines.conventional import LSQMLSolver, SimulationState, run_engine
    from phaser_amd.groups import GroupManager
    from phaser_amd.optics import make_propagators
    from phaser_amd.plan import ConventionalEnginePlan
    from phaser_amd.synthetic import CONFIG_2
    state, pats, _truth = _full(CONFIG_2)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=2, grouping=64,
                                  calc_error_fraction=1.0)
    out = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7}, plan).to_numpy()
    errs = out.progress.detector_errors
    print(f"config 2 full: detector errors {errs}", flush=True)
    assert numpy.all(numpy.isfinite(out.object.data)) and numpy.all(numpy.isfinite(out.probe.data))
    assert errs[1] < errs[0]

    dev = state.copy().to_device()
    props = make_propagators(dev.probe.sampling, dev.wavelength, dev.object.thicknesses, 2 / 3)
    groups = GroupManager(dev.scan, 64, False, seed=7)
    ws = GroupWorkspace(dev, pats.patterns, pats.pattern_mask, props, noise_offset=0.0, noise_eps=1e-3,
                        n_groups=len(groups), bmax=64, keep_chi=True)
    ws.set_positions(dev.object.sampling, dev.scan)
    solver = LSQMLSolver(plan, plan.solver, use_graphs=False)
    sim = SimulationState(state=dev, noise_model=plan.noise_model, workspace=ws)
    solver.presolve(sim, groups.iter(dev.scan), patterns=pats.patterns, pattern_mask=pats.pattern_mask,
                    propagators=props)
    obj0, probe0 = ws.object.cpu().numpy(), ws.probe.cpu().numpy()
    glist = list(groups.iter(dev.scan, 1, True))
    ws.set_groups(glist)
    solver._set_params(sim)
    ws.reset_finite_flag()
    ws.launch_lsqml_group(0, update_object=True, update_probe=True, update_position=False, calc_error=True)
    torch.cuda.synchronize()
    d_obj, d_probe = ws.object.cpu().numpy() - obj0, ws.probe.cpu().numpy() - probe0
    g_err = ws.errors[0, :64].cpu().numpy()

    o = to_oracle_state(state)
    o.probe.data = probe0.copy()
    osim = osol.Sim(o, onoise.AmplitudeNoiseModel(), numpy.float32)
    ky, kx = o.probe.sampling.recip_grid(dtype=numpy.float32)
    oprops = omake_props(ky, kx, o.probe.sampling.k_max, o.wavelength, o.object.thicknesses, 2 / 3)
    grp = glist[0]
    errors, _ = osol.lsqml_run(osim, grp, pats.patterns[tuple(grp)], pattern_mask=pats.pattern_mask, props=oprops,
                               beta_object=1.0, beta_probe=1.0, update_object=True, update_probe=True,
                               update_position=False, calc_error=True, illum_reg_object=1e-2, illum_reg_probe=1e-2,
                               gamma=1e-4)
    eo, ep = rel_err(d_obj, o.object.data - obj0), rel_err(d_probe, o.probe.data - probe0)
    ee = rel_err(g_err, numpy.asarray(errors).reshape(-1))
    print(f"config 2 full, group 0 step: object update {eo:.2e} probe update {ep:.2e} errors {ee:.2e}", flush=True)
    assert eo < 1e-4 and ep < 1e-4 and ee < 1e-5


def test_config3_full_size_iteration_and_first_group():
    """All 16384 positions of the 16-slice, 6-mode gradient-engine problem (1030^2 object,
    position correction on): one iteration is finite and a second lowers the loss; the
    gradients of group 0 (object over the group's region, probe, positions) equal the
    oracle's model_grad on that group, at the state the iterations reach."""
    from phaser_amd.engines.gradient import make_model, run_engine
    from phaser_amd.groups import GroupManager
    from phaser_amd.plan import GradientEnginePlan, NOISE_MODELS
    from phaser_amd.synthetic import SyntheticConfig
    cfg = SyntheticConfig(n=128, n_modes=6, n_slices=16, scan_shape=(128, 128), scan_step_px=6.0, jitter_px=1.0,
                          dose=5e4)
    state, pats, _truth = _full(cfg)
    plan = GradientEnginePlan(noise_model='amplitude', niter=2, grouping=64, solvers=_example_solvers(),
                              update_positions=True)
    out = run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32}, plan).to_numpy()
    losses = out.progress.detector_errors
    print(f"config 3 full: losses {losses}", flush=True)
    assert numpy.all(numpy.isfinite(out.object.data)) and numpy.all(numpy.isfinite(out.scan))
    assert losses[1] < losses[0]

    # the gradients at the state the two iterations left (a flat object has no position
    # gradient beyond rounding)
    state = out
    grp = next(iter(GroupManager(state.scan, 64, False, seed=3).iter(state.scan)))
    model, _ = make_model(state.copy(), pats.patterns, pats.pattern_mask, NOISE_MODELS['amplitude'](), groups=[grp])
    model.forward(0)
    model.backward(0)
    loss = float(model.loss_tensor(0).item())
    g_obj = model.ws.grad_obj.cpu().numpy()
    g_probe = model.ws.grad_probe.view(model.ws.probe.shape).cpu().numpy()
    g_pos = model.position_grads(0)[:64].cpu().numpy()
    st = to_oracle_state(state)
    ky, kx = st.probe.sampling.recip_grid(dtype=numpy.float32)
    from oracle.optics import make_propagators as omake_props
    oprops = omake_props(ky, kx, st.probe.sampling.k_max, st.wavelength, st.object.thicknesses, 2 / 3)
    nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
    o_loss, o_obj, o_probe, o_pos = og.model_grad(st, grp, pats.patterns, pats.pattern_mask, nz, ky, kx, oprops)
    covered = numpy.abs(o_obj) > 0
    e = (abs(loss - o_loss) / abs(o_loss), rel_err(g_obj[covered], o_obj[covered]), rel_err(g_probe, o_probe),
         rel_err(g_pos, o_pos))
    print(f"config 3 full, group 0 gradients: loss {e[0]:.2e} object {e[1]:.2e} probe {e[2]:.2e} "
          f"positions {e[3]:.2e}", flush=True)
    assert numpy.all(g_obj[~covered] == 0)
    assert e[0] < 1e-4 and e[1] < 1e-4 and e[2] < 1e-4 and e[3] < 1e-3


def test_config5_full_size_first_groups():
    """The full 24-slice problem (1540^2 x 24 slices, 192^2 x 8 modes, 36,864 positions):
    from the GPU presolve's probe, LSQ-ML groups 0 and 1 of iteration 1 (reference
    defaults, beta 1; object steps only) change the object as the oracle's lsqml_run does
    on those groups in sequence.  The steps are large (the per-slice step sizes add up over
    24 slices: after group 1 the object amplitude spans about 0.01-6), which makes the
    comparison a strong one; the values stay finite."""
    import torch
    import oracle.solvers as osol
    from oracle.optics import make_propagators as omake_props
    from phaser_amd.device import GroupWorkspace
    from phaser_amd.engines.conventional import LSQMLSolver, SimulationState
    from phaser_amd.groups import GroupManager
    from phaser_amd.optics import make_propagators
    from phaser_amd.plan import ConventionalEnginePlan
    from phaser_amd.synthetic import LSQML_192X8_S24
    state, pats, _truth = _full(LSQML_192X8_S24)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=1, grouping=64)
    dev = state.copy().to_device()
    props = make_propagators(dev.probe.sampling, dev.wavelength, dev.object.thicknesses, 2 / 3)
    groups = GroupManager(dev.scan, 64, False, seed=7)
    ws = GroupWorkspace(dev, pats.patterns, pats.pattern_mask, props, noise_offset=0.0, noise_eps=1e-3,
                        n_groups=len(groups), bmax=64, keep_chi=True)
    ws.set_positions(dev.object.sampling, dev.scan)
    solver = LSQMLSolver(plan, plan.solver, use_graphs=False)
    sim = SimulationState(state=dev, noise_model=plan.noise_model, workspace=ws)
    solver.presolve(sim, groups.iter(dev.scan), patterns=pats.patterns, pattern_mask=pats.pattern_mask,
                    propagators=props)
    obj0, probe0 = ws.object.cpu().numpy(), ws.probe.cpu().numpy()
    glist = list(groups.iter(dev.scan, 1, True))
    ws.set_groups(glist)
    solver._set_params(sim)
    ws.reset_finite_flag()

    o = to_oracle_state(state)
    o.probe.data = probe0.copy()
    osim = osol.Sim(o, onoise.AmplitudeNoiseModel(), numpy.float32)
    ky, kx = o.probe.sampling.recip_grid(dtype=numpy.float32)
    oprops = omake_props(ky, kx, o.probe.sampling.k_max, o.wavelength, o.object.thicknesses, 2 / 3)
    for g in range(2):
        print(f"config 5 full: group {g}", flush=True)
        ws.launch_lsqml_group(g, update_object=True, update_probe=False, update_position=False, calc_error=True)
        torch.cuda.synchronize()
        gobj = ws.object.cpu().numpy()
        g_err = ws.errors[g, :64].cpu().numpy()
        grp = glist[g]
        errors, _ = osol.lsqml_run(osim, grp, pats.patterns[tuple(grp)], pattern_mask=pats.pattern_mask,
                                   props=oprops, beta_object=1.0, beta_probe=1.0, update_object=True,
                                   update_probe=False, update_position=False, calc_error=True,
                                   illum_reg_object=1e-2, illum_reg_probe=1e-2, gamma=1e-4)
        eo = rel_err(gobj - obj0, o.object.data - obj0)
        ee = rel_err(g_err, numpy.asarray(errors).reshape(-1))
        amp = numpy.abs(o.object.data)
        print(f"config 5 full, group {g}: object update {eo:.2e} errors {ee:.2e}; oracle |O| in "
              f"[{amp.min():.3g}, {amp.max():.3g}]", flush=True)
        assert numpy.all(numpy.isfinite(gobj))
        assert eo < 1e-4 and ee < 1e-5
