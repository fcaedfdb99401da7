"""Per-position tilts on the GPU (tilted propagators, tilt_propagators of
phaser/engines/common/simulation.py:193-222) against the oracle on the same seeded inputs:
the conventional engines through the fused sweeps (128^2, 8 lanes per row) and the
per-slice kernels (64^2 and 192^2 planes), the gradient engine's adjoint (stored and
recomputed slices) and its engine loop.  Bars as in test_gpu_engine.py and
test_gpu_gradient.py; tilts of a few mrad, drawn per position."""
import numpy
import pytest

import oracle.gradient as og
from oracle.optics import make_propagators
from tests.helpers import rel_err, to_oracle_state
from tests.test_gpu_engine import _check, _problem, _run_all

pytestmark = pytest.mark.gpu


def _tilted(state, seed, scale=4.0):
    rng = numpy.random.default_rng(seed)
    state.tilt = rng.uniform(-scale, scale, size=state.scan.shape).astype(numpy.float32)
    return state


@pytest.mark.parametrize('solver', ['lsqml', 'epie'])
def test_conventional_tilted_n128_fused(solver):
    _, state, pats = _problem(n=128, n_modes=2, n_slices=3, scan_shape=(5, 5), scan_step_px=8.0, jitter_px=0.5,
                              seed=31, defocus=40.0)
    _tilted(state, 31)
    print(f"tilted {solver} n=128 M=2 S=3", flush=True)
    _check(*_run_all(state, pats, solver=solver, niter=2, grouping=13, use_graphs=(solver == 'lsqml')))


def test_lsqml_tilted_n64_modes_slices():
    _, state, pats = _problem(n=64, n_modes=3, n_slices=4, scan_shape=(8, 8), scan_step_px=6.0, jitter_px=0.5,
                              seed=32, defocus=30.0)
    _tilted(state, 32)
    print("tilted lsqml n=64 M=3 S=4", flush=True)
    _check(*_run_all(state, pats, solver='lsqml', niter=3, grouping=16))


def test_lsqml_tilted_n192_per_slice():
    _, state, pats = _problem(n=192, n_modes=2, n_slices=2, scan_shape=(2, 2), scan_step_px=10.0, jitter_px=0.5,
                              seed=33, defocus=30.0)
    _tilted(state, 33)
    print("tilted lsqml n=192 M=2 S=2", flush=True)
    _check(*_run_all(state, pats, solver='lsqml', niter=2, grouping=4))


def test_zero_tilts_equal_no_tilts():
    """All-zero tilts run untilted (every ramp is exactly 1): bit-identical results."""
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan
    _, state, pats = _problem(n=64, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=5.0, jitter_px=0.5,
                              seed=34, defocus=30.0)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=2, grouping=12)
    runs = []
    for tilt in (None, numpy.zeros(state.scan.shape, numpy.float32)):
        st = state.copy()
        st.tilt = tilt
        runs.append(run_engine({'data': pats, 'state': st, 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                                'recons_name': 't'}, plan).to_numpy())
    assert numpy.array_equal(runs[0].object.data, runs[1].object.data)
    assert numpy.array_equal(runs[0].probe.data, runs[1].probe.data)


@pytest.mark.parametrize('n,recompute', [(32, False), (32, True), (128, False)])
def test_tilted_adjoint_matches_oracle_gradient(n, recompute):
    from phaser_amd.engines.gradient import make_model
    from phaser_amd.plan import NOISE_MODELS
    from tests.test_gpu_gradient import _problem as gproblem
    _cfg, state, pats = gproblem(S=4, n=n)
    _tilted(state, 35 + n)
    group = numpy.array(numpy.unravel_index(numpy.arange(25), (5, 5)))
    model, _dstate = make_model(state, pats.patterns, pats.pattern_mask, NOISE_MODELS['amplitude'](),
                                groups=[group], recompute=recompute)
    model.forward(0)
    model.backward(0)
    loss = float(model.loss_tensor(0).item())
    g_obj = model.ws.grad_obj.cpu().numpy()
    g_probe = model.ws.grad_probe.view(model.ws.probe.shape).cpu().numpy()
    st = to_oracle_state(state, dtype=numpy.float64)
    ky, kx = st.probe.sampling.recip_grid(dtype=numpy.float64)
    props = make_propagators(ky, kx, st.probe.sampling.k_max, st.wavelength, st.object.thicknesses)
    nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
    o_loss, o_obj, o_probe, _ = og.model_grad(st, group, pats.patterns.astype(numpy.float64),
                                              pats.pattern_mask.astype(numpy.float64), nz, ky, kx, props)
    e = (abs(loss - o_loss) / abs(o_loss), rel_err(g_obj, o_obj), rel_err(g_probe, o_probe))
    print(f"tilted adjoint n={n} recompute={recompute}: loss {e[0]:.2e} object {e[1]:.2e} probe {e[2]:.2e}",
          flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-4


def test_gradient_engine_tilted_matches_oracle():
    from phaser_amd.engines.gradient import run_engine
    from phaser_amd.plan import GradientEnginePlan, SGDSolverPlan
    from tests.test_gpu_gradient import _problem as gproblem
    _cfg, state, pats = gproblem(S=3, M=2, scan=(6, 6), seed=36)
    state.probe.data = (state.probe.data / numpy.float32(numpy.sqrt(1e5))).astype(numpy.complex64)
    _tilted(state, 36)
    sp, hp = SGDSolverPlan(learning_rate=0.5, momentum=0.5), dict(learning_rate=0.5, momentum=0.5, nesterov=True)
    plan = GradientEnginePlan(noise_model='amplitude', solvers={('object', 'probe'): sp}, niter=3, grouping=12)
    g = run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
        og.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), noise=nz,
                      solvers=[og.Solver('sgd', ['object', 'probe'], **hp)], niter=3, grouping=12, seed=3, dtype=dt)
        res.append(o)
    o32, o64 = res
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('loss', lambda s: s.progress.detector_errors)):
        a, b, c = rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64))
        print(f"tilted sgd {name}: gpu-vs-fp32 {a:.2e} gpu-vs-fp64 {b:.2e} fp32-vs-fp64 {c:.2e}", flush=True)
        assert b <= 3 * c + 1e-4
