"""The learned initializer end to end on the GPU: a network trained on plain-LSQ-ML
reconstructions of seeded synthetic objects (3-5 slices) starts a new object closer to
the ground truth than the flat start does, and the time-to-NRMSE comparison (learned
init + adjoint gradient engine vs plain LSQ-ML from the flat start) runs to the
threshold.  NRMSE: projected object, scanned region, optimal global phase."""
import pytest

pytestmark = pytest.mark.gpu


def test_learned_init_gains_over_flat_start_and_reaches_threshold():
    from phaser_amd import learned
    mlp, info = learned.train_on_reconstructions(48, niter=40, epochs=30, lr=1e-4)
    print(info, flush=True)
    assert info['loss_first_last'][1] < info['loss_first_last'][0]
    r = learned.time_to_nrmse(learned.evaluation_config(), mlp, threshold=0.05, max_iter=60, lr_object=30.0)
    print({k: v for k, v in r.items() if k not in ('plain_lsqml', 'learned_gradient', 'flat_gradient')}, flush=True)
    for k in ('plain_lsqml', 'learned_gradient', 'flat_gradient'):
        print(k, r[k]['seconds'], r[k]['iterations'], r[k]['final_nrmse'], flush=True)
    # what the network gains over the method without it: a closer starting object
    assert r['learned_init_nrmse'] < r['flat_start_nrmse']
    assert r['plain_lsqml']['seconds'] is not None and r['learned_gradient']['seconds'] is not None
