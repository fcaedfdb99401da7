"""GPU parity of the hand-written adjoint (gradient engine) against the oracle's gradient,
which is itself pinned by finite differences and torch autograd
(tests/test_oracle_gradient.py); slice recompute vs stored waves; the
torch.autograd.Function; and the engine loop with SGD / Adam / Polyak against the oracle
gradient engine."""
import numpy
import pytest

import oracle.gradient as og
from oracle.optics import make_propagators
from tests.helpers import rel_err, to_oracle_state, oracle_simulate

pytestmark = pytest.mark.gpu


def _problem(S=3, M=2, n=32, scan=(5, 5), noise='amplitude', seed=21):
    from phaser_amd.synthetic import SyntheticConfig, make_dataset
    cfg = SyntheticConfig(n=n, n_modes=M, n_slices=S, scan_shape=scan, scan_step_px=3.0, jitter_px=0.4, seed=seed,
                          defocus=30.0)
    state, pats, truth = make_dataset(cfg, simulate=oracle_simulate)
    # a non-flat current object so every gradient term is generic
    rng = numpy.random.default_rng(seed)
    state.object.data = (state.object.data * numpy.exp(0.1j * rng.normal(size=state.object.data.shape))).astype(numpy.complex64)
    state.probe.data = (state.probe.data * numpy.float32(numpy.sqrt(1e5))).astype(numpy.complex64)
    return cfg, state, pats


def _oracle_grad(state, pats, group, noise_name):
    st = to_oracle_state(state, dtype=numpy.float64)
    ky, kx = st.probe.sampling.recip_grid(dtype=numpy.float64)
    props = make_propagators(ky, kx, st.probe.sampling.k_max, st.wavelength, st.object.thicknesses)
    nz = {'name': noise_name, 'offset': 0.375 if noise_name == 'anscombe' else 0.0, 'eps': 1e-3,
          'gaussian_variance': 0.1}
    return og.model_grad(st, group, pats.patterns.astype(numpy.float64), pats.pattern_mask.astype(numpy.float64),
                         nz, ky, kx, props)


@pytest.mark.parametrize('noise', ['amplitude', 'poisson'])
@pytest.mark.parametrize('recompute', [False, True])
def test_adjoint_matches_oracle_gradient(noise, recompute):
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import NOISE_MODELS
    _cfg, state, pats = _problem(S=4)
    group = numpy.array(numpy.unravel_index(numpy.arange(25), (5, 5)))
    model, dstate = make_model(state, pats.patterns, pats.pattern_mask, NOISE_MODELS[noise](), groups=[group],
                               recompute=recompute)
    model.forward(0)
    model.backward(0)
    loss = float(model.loss_tensor(0).item())
    g_obj = model.ws.grad_obj.cpu().numpy()
    g_probe = model.ws.grad_probe.view(model.ws.probe.shape).cpu().numpy()
    g_pos = model.position_grads(0)[:25].cpu().numpy()
    o_loss, o_obj, o_probe, o_pos = _oracle_grad(state, pats, group, noise)
    e = (abs(loss - o_loss) / abs(o_loss), rel_err(g_obj, o_obj), rel_err(g_probe, o_probe), rel_err(g_pos, o_pos))
    print(f"{noise} recompute={recompute}: loss {e[0]:.2e} object {e[1]:.2e} probe {e[2]:.2e} positions {e[3]:.2e}",
          flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-4 and e[3] < 1e-3


def test_adjoint_n128_matches_oracle_gradient():
    """The fused adjoint sweep on 128^2 planes (8 lanes per row) against the oracle."""
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import NOISE_MODELS
    _cfg, state, pats = _problem(S=3, M=2, n=128, scan=(3, 3))
    group = numpy.array(numpy.unravel_index(numpy.arange(9), (3, 3)))
    model, _ = make_model(state, pats.patterns, pats.pattern_mask, NOISE_MODELS['amplitude'](), groups=[group])
    model.forward(0)
    model.backward(0)
    loss = float(model.loss_tensor(0).item())
    o_loss, o_obj, o_probe, o_pos = _oracle_grad(state, pats, group, 'amplitude')
    e = (abs(loss - o_loss) / abs(o_loss), rel_err(model.ws.grad_obj.cpu().numpy(), o_obj),
         rel_err(model.ws.grad_probe.view(model.ws.probe.shape).cpu().numpy(), o_probe),
         rel_err(model.position_grads(0)[:9].cpu().numpy(), o_pos))
    print(f"n=128: loss {e[0]:.2e} object {e[1]:.2e} probe {e[2]:.2e} positions {e[3]:.2e}", flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-4 and e[3] < 1e-3


def test_adjoint_n192_matches_oracle_gradient():
    """The fused adjoint and the gathered object gradient on 192^2 planes."""
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import NOISE_MODELS
    _cfg, state, pats = _problem(S=2, M=2, n=192, scan=(2, 2))
    group = numpy.array(numpy.unravel_index(numpy.arange(4), (2, 2)))
    model, _ = make_model(state, pats.patterns, pats.pattern_mask, NOISE_MODELS['amplitude'](), groups=[group])
    model.forward(0)
    model.backward(0)
    loss = float(model.loss_tensor(0).item())
    o_loss, o_obj, o_probe, o_pos = _oracle_grad(state, pats, group, 'amplitude')
    e = (abs(loss - o_loss) / abs(o_loss), rel_err(model.ws.grad_obj.cpu().numpy(), o_obj),
         rel_err(model.ws.grad_probe.view(model.ws.probe.shape).cpu().numpy(), o_probe),
         rel_err(model.position_grads(0)[:4].cpu().numpy(), o_pos))
    print(f"n=192: loss {e[0]:.2e} object {e[1]:.2e} probe {e[2]:.2e} positions {e[3]:.2e}", flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-4 and e[3] < 1e-3


def test_recompute_equals_store():
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import AmplitudeNoisePlan
    _cfg, state, pats = _problem(S=6, M=2)
    group = numpy.array(numpy.unravel_index(numpy.arange(25), (5, 5)))
    res = []
    for rc in (False, True):
        model, _ = make_model(state, pats.patterns, pats.pattern_mask, AmplitudeNoisePlan(), groups=[group],
                              recompute=rc)
        model.forward(0)
        model.backward(0)
        res.append((model.ws.grad_obj.cpu().numpy(), model.ws.grad_probe.cpu().numpy(),
                    model.position_grads(0).cpu().numpy(), model.ws.psi.shape[0]))
    assert res[0][3] == 6 and res[1][3] == 3
    # store mode runs the fused adjoint sweep and gathers the object gradient per tile,
    # recompute mode the per-slice kernels with atomics: different FP32 operation orders,
    # agreeing to ~1e-6 (both paths also match the FP64 oracle above to 1e-4)
    for a, b in zip(res[0][:3], res[1][:3]):
        e = rel_err(a, b)
        print(f"store vs recompute: {e:.2e}", flush=True)
        assert e < 1e-5


@pytest.mark.parametrize('noise', ['amplitude', 'poisson'])
def test_farfield_step_in_sweep_equals_separate_launch(noise):
    """The far-field gradient step and the loss computed in the adjoint sweep's first pass
    (phx_gradient_sweep_farfield) against phx_farfield_gradient_update followed by
    phx_gradient_sweep: the same values (object gradient bit-identical; the loss and the
    probe/position gradients are summed with atomics)."""
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import AmplitudeNoisePlan, PoissonNoisePlan
    _cfg, state, pats = _problem(S=4, M=2, noise=noise)
    group = numpy.array(numpy.unravel_index(numpy.arange(25), (5, 5)))
    nplan = PoissonNoisePlan() if noise == 'poisson' else AmplitudeNoisePlan()
    res = []
    for deferred in (False, True):
        model, _ = make_model(state, pats.patterns, pats.pattern_mask, nplan, groups=[group])
        assert model.fused
        model.forward(0, loss_in_backward=deferred)
        model.backward(0)
        res.append((model.ws.grad_obj.cpu().numpy(), model.ws.grad_probe.cpu().numpy(),
                    model.position_grads(0).cpu().numpy(), float(model.loss_tensor(0))))
    numpy.testing.assert_array_equal(res[0][0], res[1][0])
    for a, b in zip(res[0][1:3], res[1][1:3]):
        assert rel_err(a, b) < 1e-6
    assert abs(res[0][3] - res[1][3]) <= 1e-9 * abs(res[0][3])


@pytest.mark.parametrize('noise,modes,sizes,tilt', [('amplitude', 6, (64, 40), False), ('poisson', 6, (64, 40), False),
                                                    ('amplitude', 7, (40, 24), False), ('amplitude', 6, (64, 40), True)])
def test_group_sweeps_as_plane_tasks_equal_two_sweeps(noise, modes, sizes, tilt):
    """phx_gradient_group_sweeps on more 128^2 planes than the MI355X's 256 CUs, as plane
    tasks: 64 positions x 6 modes = 384 planes in 3 full launches; 40 x 7 = 280 planes in
    launches of 256, 256 and 48 tasks.  A second, smaller group follows in the same slots
    (its inactive slots skipped).  Against phx_group_forward + phx_gradient_sweep_farfield:
    every plane runs the same code either way, so the stored waves, the intensity terms
    and the object and probe gradients agree to the last bit; the loss and the position
    gradients are FP64 sums with atomics.  With per-position tilts both read the group's
    tilted propagators."""
    import torch
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import NOISE_MODELS
    _cfg, state, pats = _problem(S=3 if tilt else 2, M=modes, n=128, scan=(8, 13), noise=noise, seed=29)
    if tilt:
        state.tilt = numpy.random.default_rng(31).uniform(-4.0, 4.0, size=state.scan.shape).astype(numpy.float32)
    idx = numpy.arange(8 * 13)
    b0, b1 = sizes
    groups = [numpy.array(numpy.unravel_index(idx[:b0], (8, 13))),
              numpy.array(numpy.unravel_index(idx[b0:b0 + b1], (8, 13)))]
    planes = b0 * modes
    print(f"{torch.cuda.get_device_properties(0).multi_processor_count} CUs for {planes} planes", flush=True)
    res = []
    for tasks in (False, True):
        model, _ = make_model(state, pats.patterns, pats.pattern_mask, NOISE_MODELS[noise](), groups=groups)
        assert model.fused and model.ws.B * model.ws.M == planes
        model.group_sweeps = tasks
        out = []
        for g, gs in ((0, b0), (1, b1)):
            model.forward(g, loss_in_backward=True)
            model.backward(g)
            ws = model.ws
            out.append([t.cpu().numpy() for t in (ws.psi[:, :gs], ws.intensity[:gs], ws.grad_obj, ws.grad_probe,
                                                  model.position_grads(g)[:gs])] + [float(model.loss_tensor(g))])
        res.append(out)
    for g in range(2):
        a, b = res[0][g], res[1][g]
        for name, u, v in zip(('psi', 'intensity', 'object gradient', 'probe gradient'), a[:4], b[:4]):
            assert numpy.array_equal(u, v), f"group {g}: {name} differs"
        e_pos, e_loss = rel_err(a[4], b[4]), abs(a[5] - b[5]) / abs(a[5])
        print(f"group {g}: positions {e_pos:.2e}, loss {e_loss:.2e}", flush=True)
        assert e_pos < 1e-9 and e_loss < 1e-12


def test_dry_run_from_forward_intensity_equals_per_slice():
    """The gradient engine's dry-run ratios sum(D) / sum_m |psi_{S-1} O_{S-1}|^2 from one
    fused forward launch and its stored intensity (Parseval) against the per-slice
    kernels that sum |psi O|^2 directly."""
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import AmplitudeNoisePlan
    _cfg, state, pats = _problem(S=4, M=2)
    group = numpy.array(numpy.unravel_index(numpy.arange(25), (5, 5)))
    model, _ = make_model(state, pats.patterns, pats.pattern_mask, AmplitudeNoisePlan(), groups=[group])
    assert model.fused
    out = []
    for fused in (True, False):
        model.fused = fused
        model.dry_run(0)
        out.append(model.ws.errors[0, :25].cpu().numpy().copy())
    model.fused = True
    print('dry run ratios', out[0][:4], 'rel', rel_err(out[0], out[1]), flush=True)
    assert rel_err(out[0], out[1]) < 1e-5


def test_autograd_function_backward():
    import torch
    from phaser_amd.engines.gradient import make_model, multislice_loss
    from phaser_amd.plan import AmplitudeNoisePlan
    _cfg, state, pats = _problem(S=3)
    group = numpy.array(numpy.unravel_index(numpy.arange(25), (5, 5)))
    model, dstate = make_model(state, pats.patterns, pats.pattern_mask, AmplitudeNoisePlan(), groups=[group])
    obj = dstate.object.data.requires_grad_(True)
    probe = dstate.probe.data.requires_grad_(True)
    pos = torch.zeros((25, 2), dtype=torch.float32, device='cuda', requires_grad=True)
    loss = multislice_loss(obj, probe, pos, model, 0)
    (2.0 * loss).backward()
    o_loss, o_obj, o_probe, o_pos = _oracle_grad(state, pats, group, 'amplitude')
    assert abs(float(loss) - o_loss) / o_loss < 1e-5
    assert rel_err(obj.grad.cpu().numpy(), 2 * o_obj) < 1e-4
    assert rel_err(probe.grad.cpu().numpy(), 2 * o_probe) < 1e-4
    assert rel_err(pos.grad.cpu().numpy(), 2 * o_pos) < 1e-3


@pytest.mark.parametrize('noise', ['amplitude', 'poisson'])
def test_adjoint_n256_split_passes_match_oracle_gradient(noise):
    """256^2 planes: the adjoint in split row/column launches (the far-field gradient, the
    back-propagation, the probe-gradient spectra and the position gradients at slice 0),
    the gathered object gradient and the ordered probe-gradient sum, against the
    oracle's loss and gradients."""
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import NOISE_MODELS
    _cfg, state, pats = _problem(S=2, M=2, n=256, scan=(2, 2), noise=noise)
    group = numpy.array(numpy.unravel_index(numpy.arange(4), (2, 2)))
    model, _ = make_model(state, pats.patterns, pats.pattern_mask, NOISE_MODELS[noise](), groups=[group])
    model.forward(0)
    model.backward(0)
    loss = float(model.loss_tensor(0).item())
    o_loss, o_obj, o_probe, o_pos = _oracle_grad(state, pats, group, noise)
    e = (abs(loss - o_loss) / abs(o_loss), rel_err(model.ws.grad_obj.cpu().numpy(), o_obj),
         rel_err(model.ws.grad_probe.view(model.ws.probe.shape).cpu().numpy(), o_probe),
         rel_err(model.position_grads(0)[:4].cpu().numpy(), o_pos))
    print(f"n=256 {noise}: loss {e[0]:.2e} object {e[1]:.2e} probe {e[2]:.2e} positions {e[3]:.2e}", flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-4 and e[3] < 1e-3


def test_gradient_engine_n256_matches_oracle():
    """The gradient engine (momentum SGD on object and probe) on 256^2 planes, three
    iterations against the oracle's engine."""
    from phaser_amd.engines.gradient import run_engine
    from phaser_amd.plan import GradientEnginePlan, SGDSolverPlan
    _cfg, state, pats = _problem(S=2, M=2, n=256, scan=(3, 3), seed=24)
    state.probe.data = (state.probe.data / numpy.float32(numpy.sqrt(1e5))).astype(numpy.complex64)
    plan = GradientEnginePlan(noise_model='amplitude',
                              solvers={('object', 'probe'): SGDSolverPlan(learning_rate=0.5, momentum=0.5)},
                              niter=3, grouping=9)
    g = run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
        og.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), noise=nz,
                      solvers=[og.Solver('sgd', ['object', 'probe'], learning_rate=0.5, momentum=0.5, nesterov=True)],
                      niter=3, grouping=9, seed=3, dtype=dt)
        res.append(o)
    o32, o64 = res
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('loss', lambda s: s.progress.detector_errors)):
        a, b, c = rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64))
        print(f"n=256 {name}: gpu-vs-fp32 {a:.2e} gpu-vs-fp64 {b:.2e} fp32-vs-fp64 {c:.2e}", flush=True)
        assert b <= 3 * c + 1e-4


@pytest.mark.parametrize('solver', ['sgd', 'adam', 'polyak_sgd'])
def test_gradient_engine_matches_oracle(solver):
    import oracle.state as ost
    from phaser_amd.engines.gradient import run_engine
    from phaser_amd.plan import GradientEnginePlan, SGDSolverPlan, AdamSolverPlan, PolyakSGDSolverPlan
    _cfg, state, pats = _problem(S=2, M=2, scan=(6, 6), seed=23)
    state.probe.data = (state.probe.data / numpy.float32(numpy.sqrt(1e5))).astype(numpy.complex64)
    if solver == 'sgd':
        sp, hp = SGDSolverPlan(learning_rate=0.5, momentum=0.5), dict(learning_rate=0.5, momentum=0.5, nesterov=True)
    elif solver == 'adam':
        sp, hp = AdamSolverPlan(learning_rate=1e-3), dict(learning_rate=1e-3)
    else:
        sp, hp = PolyakSGDSolverPlan(max_learning_rate=1.0, f_min=0.0, scaling=0.5), \
            dict(max_learning_rate=1.0, f_min=0.0, scaling=0.5)
    plan = GradientEnginePlan(noise_model='amplitude', solvers={('object', 'probe'): sp}, niter=3, grouping=12)
    g = run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
        og.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), noise=nz,
                      solvers=[og.Solver(solver, ['object', 'probe'], **hp)], niter=3, grouping=12, seed=3, dtype=dt)
        res.append(o)
    o32, o64 = res
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('loss', lambda s: s.progress.detector_errors)):
        a, b, c = rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64))
        print(f"{solver} {name}: gpu-vs-fp32 {a:.2e} gpu-vs-fp64 {b:.2e} fp32-vs-fp64 {c:.2e}", flush=True)
        assert b <= 3 * c + 1e-4


def test_gather_clearing_next_group_matches_separate_clear(monkeypatch):
    """The object-gradient gather that also zeroes the next group's accumulators
    (phx_gradient_object_gather_begin) against phx_group_begin before every group (SGD
    with momentum, 3 groups, 3 iterations).  Every sum that feeds the object and the
    probe is ordered (object-gradient gather, per-position probe spectra summed in
    position order, the four modes' intensities in mode order), so two runs agree to
    the last bit and the fold must not change a bit either; the loss is an FP64 sum of
    per-position losses added in any order, read as FP32."""
    import phaser_amd.engines.gradient as gr
    from phaser_amd.plan import GradientEnginePlan, SGDSolverPlan
    _cfg, state, pats = _problem(S=2, M=4, scan=(6, 6), seed=23)
    state.probe.data = (state.probe.data / numpy.float32(numpy.sqrt(1e5))).astype(numpy.complex64)
    plan = GradientEnginePlan(noise_model='amplitude',
                              solvers={('object', 'probe'): SGDSolverPlan(learning_rate=0.5, momentum=0.5)},
                              niter=3, grouping=12)
    runs = []
    for fold in (True, True, False):
        monkeypatch.setattr(gr, 'FOLD_GROUP_BEGIN', fold)
        runs.append(gr.run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32},
                                  plan).to_numpy())
    a, a2, b = runs
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('loss', lambda s: s.progress.detector_errors)):
        noise, e = rel_err(f(a), f(a2)), rel_err(f(a), f(b))
        print(f"{name}: run-to-run {noise:.2e}, fold vs separate {e:.2e}", flush=True)
        bar = 1e-6 if name == 'loss' else 0.0
        assert noise <= bar
        assert e <= bar
