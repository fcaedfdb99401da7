"""Pins the oracle's gradient of the gradient-engine loss (oracle/gradient.py):
  * central finite differences (step 1e-3) of the FP64 loss for object, probe and
    positions, amplitude and Poisson models;
  * torch's own autograd (CPU, float64) through an independent torch.fft implementation
    of the same forward model (torch's gradient convention, 2 dL/dconj(z)).
jax is absent, so the comparison with the jax engine's gradient is unpinned."""
import numpy
import pytest

import oracle.gradient as og
from oracle.optics import make_propagators
from tests.helpers import rel_err, to_oracle_state


def _small(noise='amplitude', S=3, M=2, seed=11):
    from phaser_amd.synthetic import SyntheticConfig, make_ground_truth, thicknesses, initial_state
    from tests.helpers import oracle_simulate
    cfg = SyntheticConfig(n=16, n_modes=M, n_slices=S, scan_shape=(3, 3), scan_step_px=2.0, jitter_px=0.4,
                          seed=seed, defocus=10.0, dose=1e3)
    truth, osamp, scan, probe, wl, psamp = make_ground_truth(cfg)
    pats = oracle_simulate(truth, osamp, scan, probe, wl, psamp, thicknesses(cfg)) * cfg.dose
    pats = numpy.random.default_rng(seed).poisson(pats).astype(numpy.float64)
    state = initial_state(cfg, osamp, scan, probe, wl, psamp)
    st = to_oracle_state(state, dtype=numpy.float64)
    # perturb the flat start so the gradient is generic
    rng = numpy.random.default_rng(seed + 1)
    st.object.data = st.object.data * numpy.exp(0.2j * rng.normal(size=st.object.data.shape))
    st.probe.data = st.probe.data * 30.0
    ky, kx = st.probe.sampling.recip_grid(dtype=numpy.float64)
    props = make_propagators(ky, kx, st.probe.sampling.k_max, st.wavelength, st.object.thicknesses)
    nz = {'name': noise, 'offset': 0.375 if noise == 'anscombe' else 0.0, 'eps': 1e-3, 'gaussian_variance': 0.1}
    group = numpy.array(numpy.unravel_index(numpy.arange(9), (3, 3)))
    return st, pats, numpy.ones((16, 16)), nz, ky, kx, props, group


def _fd_complex(f, arr, idx, h=1e-3):
    base = arr[idx]
    arr[idx] = base + h
    lp = f()
    arr[idx] = base - h
    lm = f()
    arr[idx] = base + 1j * h
    ip = f()
    arr[idx] = base - 1j * h
    im = f()
    arr[idx] = base
    return (lp - lm) / (2 * h) + 1j * (ip - im) / (2 * h)


@pytest.mark.parametrize('noise', ['amplitude', 'anscombe', 'poisson'])
def test_gradient_matches_finite_differences(noise):
    st, pats, mask, nz, ky, kx, props, group = _small(noise)
    f = lambda: og.model_loss(st, group, pats, mask, nz, ky, kx, props)  # noqa: E731
    loss, g_obj, g_probe, g_pos = og.model_grad(st, group, pats, mask, nz, ky, kx, props)
    assert loss == pytest.approx(f(), rel=1e-12)
    rng = numpy.random.default_rng(0)
    S, H, W = st.object.data.shape
    # object pixels inside the scanned region, every slice
    obj_idx = [(s, int(y), int(x)) for s in range(S) for y, x in rng.integers(H // 2 - 4, H // 2 + 4, size=(3, 2))]
    fd = numpy.array([_fd_complex(f, st.object.data, i) for i in obj_idx])
    an = numpy.array([g_obj[i] for i in obj_idx])
    assert rel_err(an, fd) < 1e-6
    pr_idx = [(m, int(y), int(x)) for m in range(st.probe.data.shape[0]) for y, x in rng.integers(4, 12, size=(3, 2))]
    fd = numpy.array([_fd_complex(f, st.probe.data, i) for i in pr_idx])
    an = numpy.array([g_probe[i] for i in pr_idx])
    assert rel_err(an, fd) < 1e-6
    # positions (real): central differences of the scan
    # (entries whose +-h step moves the rounded cutout start are discontinuous; skip them)
    fd = numpy.zeros((9, 2))
    keep = numpy.ones((9, 2), dtype=bool)
    for k in range(9):
        for d in range(2):
            iy, ix = numpy.unravel_index(k, (3, 3))
            base = st.scan[iy, ix, d]
            st.scan[iy, ix, d] = base + 1e-4
            s_p = st.object.sampling.start_idxs(st.scan[iy, ix], (16, 16))
            lp = f()
            st.scan[iy, ix, d] = base - 1e-4
            s_m = st.object.sampling.start_idxs(st.scan[iy, ix], (16, 16))
            lm = f()
            st.scan[iy, ix, d] = base
            keep[k, d] = numpy.array_equal(s_p, s_m)
            fd[k, d] = (lp - lm) / 2e-4
    assert keep.sum() >= 12
    assert rel_err(g_pos[keep], fd[keep]) < 1e-5


def test_gradient_matches_torch_autograd_cpu():
    torch = pytest.importorskip("torch")
    st, pats, mask, nz, ky, kx, props, group = _small('amplitude')
    loss, g_obj, g_probe, _g_pos = og.model_grad(st, group, pats, mask, nz, ky, kx, props)

    osamp = st.object.sampling
    n = 16
    scan = st.scan[tuple(group)]
    starts = osamp.start_idxs(scan, (n, n))
    shifts = osamp.get_subpx_shifts(scan, (n, n))
    obj = torch.tensor(st.object.data, requires_grad=True)
    probe = torch.tensor(st.probe.data, requires_grad=True)
    tky, tkx = torch.tensor(ky), torch.tensor(kx)

    def fft2(a):
        return torch.fft.fft2(torch.fft.ifftshift(a, dim=(-2, -1)), norm='ortho')

    def ifft2(a):
        return torch.fft.fftshift(torch.fft.ifft2(a, norm='ortho'), dim=(-2, -1))

    cut = torch.stack([obj[:, y:y + n, x:x + n] for (y, x) in starts])          # (B, S, n, n)
    sh = torch.tensor(shifts)
    filt = torch.exp(-2j * numpy.pi * (sh[:, 1, None, None] * tkx + sh[:, 0, None, None] * tky))[:, None]
    psi = ifft2(fft2(probe)[None] * filt)
    tprops = torch.tensor(props)
    for s in range(obj.shape[0] - 1):
        psi = ifft2(fft2(psi * cut[:, s, None]) * tprops[s])
    Psi = fft2(psi * cut[:, -1, None])
    I = torch.sum(Psi.real**2 + Psi.imag**2, dim=1)
    D = torch.clamp(torch.tensor(pats[tuple(group)]), min=0.0)
    tl = 2.0 * torch.sum(torch.tensor(mask) * (torch.sqrt(D) - torch.sqrt(I) - 1e-3)**2) / 1.1
    tl.backward()
    assert float(tl) == pytest.approx(loss, rel=1e-12)
    assert rel_err(g_obj, obj.grad.numpy()) < 1e-10
    assert rel_err(g_probe, probe.grad.numpy()) < 1e-10


def test_optimizer_restatements_closed_forms():
    """optax update equations on scalars: SGD+nesterov trace, Adam bias correction,
    Polyak step clipping."""
    s = og.Solver('sgd', ['object'], learning_rate=0.5, momentum=0.9, nesterov=True)
    u1 = s.update({'object': numpy.array([1.0])}, 0.0)['object']
    u2 = s.update({'object': numpy.array([1.0])}, 0.0)['object']
    assert u1[0] == pytest.approx(0.5 * (1 + 0.9 * 1))          # t=1, u = d + m t
    assert u2[0] == pytest.approx(0.5 * (1 + 0.9 * 1.9))        # t=1.9
    a = og.Solver('adam', ['object'], learning_rate=1.0)
    ua = a.update({'object': numpy.array([2.0])}, 0.0)['object']
    assert ua[0] == pytest.approx(2.0 / (2.0 + 1e-8))           # first Adam step = sign * lr
    p = og.Solver('polyak_sgd', ['object'], scaling=1.0, max_learning_rate=10.0, f_min=0.0)
    up = p.update({'object': numpy.array([3.0, 4.0])}, 5.0)['object']
    numpy.testing.assert_allclose(up, numpy.array([3.0, 4.0]) * (5.0 / 25.0))


@pytest.mark.parametrize('nesterov', [False, True])
def test_adam_three_steps_against_explicit_sums(nesterov):
    """optax scale_by_adam (gradient/solvers.py:117-127) over three steps with b1 != b2,
    eps_root != 0 and complex directions, against the moments written as explicit sums
    m_t = (1 - b1) sum_k b1^(t-k) g_k, v_t = (1 - b2) sum_k b2^(t-k) |g_k|^2, the bias
    corrections 1 - b^t and (nesterov) m_hat = b1 m_t/(1 - b1^(t+1)) + (1 - b1) g_t/(1 - b1^t)."""
    b1, b2, eps, eps_root, lr = 0.8, 0.95, 1e-3, 0.25, 0.3
    rng = numpy.random.default_rng(2)
    gs = [rng.normal(size=4) + 1j * rng.normal(size=4) for _ in range(3)]
    a = og.Solver('adam', ['object'], learning_rate=lr, b1=b1, b2=b2, eps=eps, eps_root=eps_root, nesterov=nesterov)
    for t in range(1, 4):
        got = a.update({'object': gs[t - 1]}, 0.0)['object']
        m = (1 - b1) * sum(b1**(t - k) * gs[k - 1] for k in range(1, t + 1))
        v = (1 - b2) * sum(b2**(t - k) * numpy.abs(gs[k - 1])**2 for k in range(1, t + 1))
        if nesterov:
            mh = b1 * m / (1 - b1**(t + 1)) + (1 - b1) * gs[t - 1] / (1 - b1**t)
        else:
            mh = m / (1 - b1**t)
        want = lr * mh / (numpy.sqrt(v / (1 - b2**t) + eps_root) + eps)
        numpy.testing.assert_allclose(got, want, rtol=1e-12)
    # a constant direction: m_hat = g, v_hat = |g|^2 at every step (no nesterov)
    c = og.Solver('adam', ['object'], learning_rate=1.0, b1=b1, b2=b2, eps=0.0, eps_root=0.0)
    for _ in range(3):
        u = c.update({'object': numpy.array([3 - 4j])}, 0.0)['object']
        numpy.testing.assert_allclose(u, [(3 - 4j) / 5], rtol=1e-12)


def test_gradient_engine_step_direction_and_scale_by_finite_differences():
    """One iteration, one group of all B positions, plain SGD with lr = 1: the object moves
    by -g/(probe_int B) and (object frozen) the probe by -g/B, g = dL/dRe + i dL/dIm from
    central finite differences of the FP64 loss at the rescaled state (gradient/run.py:
    324-330 with jax's -conj(g_jax) = -g).  Independent of model_grad."""
    import copy
    state, pats, mask, nz, ky, kx, props, group = _small()
    B = group.shape[-1]
    rng = numpy.random.default_rng(4)
    S, H, W = state.object.data.shape
    for var in ('object', 'probe'):
        st = copy.deepcopy(state)
        before = copy.deepcopy(state)
        og.run_engine(st, pats, mask, noise=nz, solvers=[og.Solver('sgd', [var], learning_rate=1.0)], niter=1,
                      grouping=B, seed=1, dtype=numpy.float64, update_object=(var == 'object'),
                      update_probe=(var == 'probe'))
        # the engine's dry-run rescale (gradient/run.py:196-209)
        factors = og.dry_run(before, group, pats, ky, kx, props)
        before.probe.data = before.probe.data * numpy.sqrt(numpy.mean(factors))
        probe_int = numpy.sum(numpy.abs(before.probe.data)**2)
        arr_b = before.object.data if var == 'object' else before.probe.data
        arr_a = st.object.data if var == 'object' else st.probe.data
        scale = probe_int * B if var == 'object' else B
        f = lambda: og.model_loss(before, group, pats, mask, nz, ky, kx, props)  # noqa: E731
        for _ in range(6):
            if var == 'object':
                idx = (int(rng.integers(0, S)), int(rng.integers(H // 2 - 4, H // 2 + 4)),
                       int(rng.integers(W // 2 - 4, W // 2 + 4)))
            else:
                idx = tuple(int(rng.integers(0, d)) for d in arr_b.shape)
            want = -_fd_complex(f, arr_b, idx, h=1e-6) / scale
            got = arr_a[idx] - arr_b[idx]
            assert abs(got - want) <= 1e-5 * abs(want) + 1e-14, (var, idx, got, want)
