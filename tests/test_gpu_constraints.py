"""Constraints and cost regularizers on the GPU (phaser_amd/csrc/phx_constraints.hip,
probe_support_k) against the oracle's restatement of phaser/engines/common/
regularizers.py: each kernel on random states, then the engines with constraints and
regularizers in their plans against the oracle engines (relative L2 errors)."""
import ctypes

import numpy
import pytest

import oracle.engine as oeng
import oracle.gradient as og
import oracle.noise_models as onoise
import oracle.regularizers as oreg
from tests.helpers import oracle_simulate, rel_err, to_oracle_state

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _obj(S, H, W, seed, amp=0.3):
    rng = numpy.random.default_rng(seed)
    return (rng.normal(1, amp, (S, H, W)) + 1j * rng.normal(0, amp, (S, H, W))).astype(numpy.complex64)


def _ns_state(obj, sampling=(0.2, 0.2), thick=None, scan_shape=(6, 7)):
    from types import SimpleNamespace as NS
    S = obj.shape[0]
    return NS(object=NS(data=obj, thicknesses=numpy.full(S, 20.0) if thick is None else thick,
                        sampling=NS(sampling=numpy.asarray(sampling))),
              scan=numpy.zeros((*scan_shape, 2), numpy.float32))


def test_clamp_kernel_matches_oracle():
    import torch
    from phaser_amd import _lib
    from phaser_amd.device import stream_handle
    obj = _obj(3, 50, 61, 1)
    obj[1, 2, 3] = 0
    t = torch.from_numpy(obj).cuda()
    _lib.call("phx_clamp_amplitude_c64", _p(t), t.numel(), 1.05, None, 0, stream_handle())
    st = _ns_state(obj.copy())
    oreg.ClampObjectAmplitude(1.05).apply_iter(st, None)
    got = t.cpu().numpy()
    assert numpy.isnan(got[1, 2, 3]) and numpy.isnan(st.object.data[1, 2, 3])
    m = ~numpy.isnan(st.object.data)
    assert rel_err(got[m], st.object.data[m]) < 1e-6
    assert numpy.abs(got[m]).max() <= 1.05 * (1 + 1e-6)


@pytest.mark.parametrize('S,sigma', [(6, 30.0), (24, 50.0), (2, 50.0)])
def test_layers_kernel_matches_oracle(S, sigma):
    import torch
    from phaser_amd.constraints import RegularizeLayers
    from phaser_amd.plan import RegularizeLayersPlan
    from types import SimpleNamespace as NS
    obj = _obj(S, 37, 45, S)
    t = torch.from_numpy(obj).cuda()
    ws = NS(object=t, S=S, H=37, W=45, device='cuda')
    plan = RegularizeLayersPlan(sigma=sigma, weight=0.9)
    RegularizeLayers(plan).init_state(ws, _ns_state(obj)).apply(ws)
    st = _ns_state(obj.copy())
    oreg.RegularizeLayers(sigma=sigma, weight=0.9).apply_iter(st, None)
    e = rel_err(t.cpu().numpy(), st.object.data)
    print(f"layers S={S}: {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize('shape', [(2, 37, 41), (1, 518, 518), (3, 130, 1030)])
@pytest.mark.parametrize('kind', ['low_pass', 'gaussian'])
def test_object_filter_kernel_matches_oracle(shape, kind):
    """Bluestein line transforms at odd, composite and 1030-wide sides against numpy's
    FFT path of the reference."""
    import torch
    from phaser_amd.constraints import ObjGaussian, ObjLowPass
    from phaser_amd.plan import ObjGaussianPlan, ObjLowPassPlan
    from types import SimpleNamespace as NS
    S, H, W = shape
    obj = _obj(S, H, W, H + W)
    t = torch.from_numpy(obj).cuda()
    ws = NS(object=t, S=S, H=H, W=W, device='cuda')
    st = _ns_state(obj.copy(), sampling=(0.2, 0.25))
    if kind == 'low_pass':
        ObjLowPass(ObjLowPassPlan(max_freq=0.3)).init_state(ws, st).apply(ws)
        o = oreg.ObjLowPass(max_freq=0.3)
    else:
        ObjGaussian(ObjGaussianPlan(sigma=0.6, weight=0.8)).init_state(ws, st).apply(ws)
        o = oreg.ObjGaussian(sigma=0.6, weight=0.8)
    o.apply_iter(st, o.init_state(st))
    e = rel_err(t.cpu().numpy(), st.object.data)
    print(f"{kind} {shape}: {e:.2e}", flush=True)
    assert e < 1e-5


def test_cost_regularizer_kernel_matches_oracle():
    import torch
    from phaser_amd.constraints import CostRegularizers
    from phaser_amd.plan import CostRegularizerPlan
    from types import SimpleNamespace as NS
    S, H, W = 4, 33, 40
    obj = _obj(S, H, W, 7)
    t = torch.from_numpy(obj).cuda()
    g0 = _obj(S, H, W, 8)
    grad = torch.from_numpy(g0).cuda()
    loss = torch.zeros(1, dtype=torch.float64, device='cuda')
    ws = NS(object=t, grad_obj=grad, S=S, H=H, W=W, P=42)
    costs = {'obj_l1': 0.3, 'obj_l2': 1.0, 'obj_tikh': 5.0, 'layers_tv': 0.2, 'layers_tikh': 0.7}
    regs = CostRegularizers([CostRegularizerPlan(cost=c, name=n) for n, c in costs.items()], S)
    regs.apply(ws, 12, _p(loss), True)
    st = _ns_state(obj.astype(numpy.complex128))
    group = numpy.zeros((2, 12), numpy.int64)
    o_loss, o_grad = 0.0, g0.astype(numpy.complex128)
    for n, c in costs.items():
        r = oreg.REGULARIZERS[n](cost=c)
        o_loss += r.calc_loss_group(group, st, None)[0]
        o_grad = o_grad + r.grad_group(group, st)
    print(f"loss {float(loss.item()):.6e} vs {o_loss:.6e}", flush=True)
    assert abs(float(loss.item()) - o_loss) / o_loss < 1e-5
    assert rel_err(grad.cpu().numpy(), o_grad) < 1e-5


def test_probe_support_kernel_matches_oracle():
    """Through a group workspace (the plane kernels' descriptor), 64^2 and 128^2 planes;
    probe_k is left as fft2 of the new probe."""
    import torch
    from phaser_amd import _lib
    from phaser_amd.constraints import LimitProbeSupport
    from phaser_amd.device import stream_handle
    from phaser_amd.plan import LimitProbeSupportPlan
    from oracle.num import Sampling
    from types import SimpleNamespace as NS
    for n in (64, 128):
        ws, state = _workspace(n=n, M=3)
        P0 = ws.probe.cpu().numpy()
        con = LimitProbeSupport(LimitProbeSupportPlan(max_angle=18.0)).init_state(ws, state)
        con.apply(ws, 0)
        got = ws.probe.cpu().numpy()
        ost = NS(probe=NS(data=P0.astype(numpy.complex128), sampling=Sampling((n, n), sampling=tuple(
            state.probe.sampling.sampling))), wavelength=state.wavelength)
        o = oreg.LimitProbeSupport(max_angle=18.0)
        o.apply_iter(ost, o.init_state(ost))
        e = rel_err(got, ost.probe.data)
        # probe_k against the spectra phx_probe_spectra computes for the new probe
        pk = ws.probe_k.clone()
        _lib.call("phx_probe_spectra", ws.desc_refs[0], stream_handle())
        torch.cuda.synchronize()
        ek = rel_err(pk.cpu().numpy(), ws.probe_k.cpu().numpy())
        print(f"n={n}: probe {e:.2e}, spectra {ek:.2e}", flush=True)
        assert e < 1e-5 and ek < 1e-6


def _workspace(n, M):
    from phaser_amd.device import GroupWorkspace
    from phaser_amd.groups import GroupManager
    from phaser_amd.optics import make_propagators
    from phaser_amd.synthetic import SyntheticConfig, make_dataset
    cfg = SyntheticConfig(n=n, n_modes=M, n_slices=2, scan_shape=(3, 3), scan_step_px=6.0, jitter_px=0.5, seed=n)
    state, pats, _ = make_dataset(cfg, simulate=oracle_simulate)
    state = state.to_device()
    props = make_propagators(state.probe.sampling, state.wavelength, state.object.thicknesses, 2 / 3)
    groups = list(GroupManager(state.scan, 9, False, seed=3).iter(state.scan))
    ws = GroupWorkspace(state, pats.patterns, pats.pattern_mask, props, noise_offset=0.0, noise_eps=1e-3,
                        n_groups=len(groups), bmax=9, keep_chi=True)
    ws.set_positions(state.object.sampling, state.scan)
    ws.set_groups(groups)
    return ws, state


# ------------------------------------------------------------------ engines
def _problem(**kw):
    from phaser_amd.synthetic import SyntheticConfig, make_dataset
    state, pats, _truth = make_dataset(SyntheticConfig(**kw), simulate=oracle_simulate)
    return state, pats


def _oracle_cons(specs):
    return [oreg.make(s, oreg.CONSTRAINTS) for s in specs]


@pytest.mark.parametrize('solver', ['lsqml', 'epie'])
def test_conventional_engine_with_constraints(solver):
    """Group constraints (clamp, low-pass) and iter constraints (layers, probe support,
    Gaussian) in a 3-slice, 2-mode run with graphs, against the oracle engine."""
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, SOLVERS
    state, pats = _problem(n=64, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=5.0, jitter_px=0.5, seed=31)
    gcs = [{'type': 'clamp_object_amplitude', 'amplitude': 1.0}, {'type': 'obj_low_pass', 'max_freq': 0.45}]
    ics = [{'type': 'layers', 'sigma': 30.0, 'weight': 0.5}, {'type': 'limit_probe_support', 'max_angle': 22.0},
           {'type': 'obj_gaussian', 'sigma': 0.3, 'weight': 0.5}]
    plan = ConventionalEnginePlan(noise_model='amplitude', solver=SOLVERS[solver](beta_object=0.5, beta_probe=0.5),
                                  niter=3, grouping=12, group_constraints=gcs, iter_constraints=ics)
    g = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7, 'observer': None,
                    'recons_name': 't'}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        oeng.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), dtype=dt, solver=solver,
                        noise_model=onoise.AmplitudeNoiseModel(), niter=3, grouping=12, seed=7, beta_object=0.5,
                        beta_probe=0.5, group_constraints=_oracle_cons(gcs), iter_constraints=_oracle_cons(ics))
        res.append(o)
    o32, o64 = res
    assert numpy.abs(g.object.data).max() <= 1.0 + 1e-3     # clamp then low-pass
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('errors', lambda s: s.progress.detector_errors)):
        a, b, c = rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64))
        print(f"{solver} {name}: gpu-vs-fp32 {a:.2e} gpu-vs-fp64 {b:.2e} fp32-vs-fp64 {c:.2e}", flush=True)
        assert a < 1e-4 if c < 1e-5 else b <= 3 * c + 1e-5


def test_gradient_engine_with_regularizers_and_constraints():
    """Cost regularizers (obj_l2, obj_tikh, layers_tikh, as the example plans use them),
    a clamp group constraint and probe-support + clamp iter constraints, SGD on object
    and probe, against the oracle gradient engine."""
    from phaser_amd.engines.gradient import run_engine
    from phaser_amd.plan import GradientEnginePlan, SGDSolverPlan
    state, pats = _problem(n=32, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5, seed=23)
    state.probe.data = (state.probe.data / numpy.float32(numpy.sqrt(1e5))).astype(numpy.complex64)
    regs = [{'type': 'obj_l2', 'cost': 0.4}, {'type': 'obj_tikh', 'cost': 0.2}, {'type': 'layers_tikh', 'cost': 0.3}]
    gcs = [{'type': 'clamp_object_amplitude', 'amplitude': 1.02}]
    ics = [{'type': 'limit_probe_support', 'max_angle': 22.0}, {'type': 'clamp_object_amplitude', 'amplitude': 1.0}]
    plan = GradientEnginePlan(noise_model='amplitude', niter=3, grouping=12, regularizers=regs,
                              group_constraints=gcs, iter_constraints=ics,
                              solvers={('object', 'probe'): SGDSolverPlan(learning_rate=0.5, momentum=0.5)})
    g = run_engine({'data': pats, 'state': state.copy(), 'seed': 3, 'dtype': numpy.float32}, plan).to_numpy()
    res = []
    for dt in (numpy.float32, numpy.float64):
        o = to_oracle_state(state, dtype=dt)
        nz = {'name': 'amplitude', 'offset': 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
        og.run_engine(o, pats.patterns.astype(dt), pats.pattern_mask.astype(dt), noise=nz, dtype=dt,
                      solvers=[og.Solver('sgd', ['object', 'probe'], learning_rate=0.5, momentum=0.5, nesterov=True)],
                      niter=3, grouping=12, seed=3,
                      regularizers=[oreg.make(r, oreg.REGULARIZERS) for r in regs],
                      group_constraints=_oracle_cons(gcs), iter_constraints=_oracle_cons(ics))
        res.append(o)
    o32, o64 = res
    for name, f in (('object', lambda s: s.object.data), ('probe', lambda s: s.probe.data),
                    ('loss', lambda s: s.progress.detector_errors)):
        a, b, c = rel_err(f(g), f(o32)), rel_err(f(g), f(o64)), rel_err(f(o32), f(o64))
        print(f"gradient {name}: gpu-vs-fp32 {a:.2e} gpu-vs-fp64 {b:.2e} fp32-vs-fp64 {c:.2e}", flush=True)
        assert a < 1e-4 if c < 1e-5 else b <= 3 * c + 1e-5


def test_unsupported_constraint_raises_before_running():
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan
    state, pats = _problem(n=32, n_modes=1, n_slices=1, scan_shape=(3, 3), scan_step_px=4.0, jitter_px=0.5, seed=1)
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=1,
                                  iter_constraints=[{'type': 'remove_phase_ramp'}])
    with pytest.raises(NotImplementedError, match='remove_phase_ramp'):
        run_engine({'data': pats, 'state': state.copy(), 'seed': 1}, plan)
