"""GPU parity of the conventional engines against the CPU oracle on the same seeded
inputs: object, probe and per-iteration detector error in complex64 must agree within
1e-4 relative (L2 norm of the difference over the L2 norm of the oracle result)."""
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


def _run_both(state, pats, *, solver, niter, grouping, use_graphs=False, position=None, **plan_kw):
    from phaser_amd.engines.conventional import run_engine
    from phaser_amd.plan import ConventionalEnginePlan, MomentumPositionSolverPlan
    plan = ConventionalEnginePlan(noise_model=plan_kw.pop('noise', 'amplitude'), solver=solver, niter=niter,
                                  grouping=grouping,
                                  position_solver=MomentumPositionSolverPlan(**position) if position else None,
                                  update_positions=bool(position), **plan_kw)
    ostate = to_oracle_state(state)
    gstate = run_engine({'data': pats, 'state': state.copy(), 'dtype': numpy.float32, 'seed': 7,
                         'observer': None, 'recons_name': 't'}, plan, use_graphs=use_graphs).to_numpy()
    noise = (onoise.AnscombeNoiseModel() if plan.noise_model.name == 'anscombe' else onoise.AmplitudeNoiseModel())
    opos = oeng.MomentumPositionSolver(**position) if position else None
    oeng.run_engine(ostate, pats.patterns, pats.pattern_mask, solver=solver, noise_model=noise, niter=niter,
                    grouping=grouping, seed=7, update_positions=bool(position), position_solver=opos,
                    calc_error_fraction=plan.calc_error_fraction,
                    update_probe=plan.update_probe, update_object=plan.update_object)
    return gstate, ostate


def _check(gstate, ostate, tol=TOL):
    eo = rel_err(gstate.object.data, ostate.object.data)
    ep = rel_err(gstate.probe.data, ostate.probe.data)
    ee = rel_err(gstate.progress.detector_errors, ostate.progress.detector_errors)
    print(f"rel err object {eo:.2e} probe {ep:.2e} errors {ee:.2e}", flush=True)
    assert len(gstate.progress.detector_errors) == len(ostate.progress.detector_errors)
    assert eo < tol and ep < tol and ee < tol
    return eo, ep, ee


def test_lsqml_one_iteration_multislice_small():
    cfg, state, pats = _problem(n=32, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5, seed=1)
    g, o = _run_both(state, pats, solver='lsqml', niter=1, grouping=12)
    _check(g, o)


def test_epie_one_iteration_multislice_small():
    cfg, state, pats = _problem(n=32, n_modes=2, n_slices=3, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5, seed=2)
    g, o = _run_both(state, pats, solver='epie', niter=1, grouping=12)
    _check(g, o)


def test_lsqml_single_slice_single_mode():
    cfg, state, pats = _problem(n=32, n_modes=1, n_slices=1, scan_shape=(6, 6), scan_step_px=3.0, jitter_px=0.5, seed=3)
    g, o = _run_both(state, pats, solver='lsqml', niter=3, grouping=12)
    _check(g, o)


def test_epie_config1_ten_iterations():
    """Config 1: single-slice ePIE, 64^2 probe, 16x16 jittered scan, 10 iterations."""
    cfg, state, pats = _problem(n=64, n_modes=1, n_slices=1, scan_shape=(16, 16), scan_step_px=4.0,
                                jitter_px=0.5, seed=0)
    g, o = _run_both(state, pats, solver='epie', niter=10, grouping=64)
    _check(g, o)


def test_lsqml_multislice_modes_five_iterations_with_graphs():
    """Reduced config 2 (LSQ-ML, 4 modes, 4 slices, jitter-free raster with .5 index ties),
    whole iterations captured in a HIP graph and replayed."""
    cfg, state, pats = _problem(n=64, n_modes=4, n_slices=4, scan_shape=(12, 12), scan_step_px=6.0,
                                jitter_px=0.0, seed=4)
    g, o = _run_both(state, pats, solver='lsqml', niter=5, grouping=48, use_graphs=True)
    _check(g, o)


def test_lsqml_position_correction():
    cfg, state, pats = _problem(n=32, n_modes=2, n_slices=2, scan_shape=(8, 8), scan_step_px=3.0, jitter_px=0.5, seed=5)
    g, o = _run_both(state, pats, solver='lsqml', niter=3, grouping=16,
                     position={'step_size': 1e-2, 'max_step_size': 0.1, 'momentum': 0.9})
    _check(g, o)
    ds = rel_err(g.scan - state.scan, o.scan - state.scan)
    print(f"scan update rel err {ds:.2e}", flush=True)
    assert ds < 1e-3
