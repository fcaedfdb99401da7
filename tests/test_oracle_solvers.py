"""Pins of the oracle's LSQ-ML / ePIE group steps (oracle/solvers.py), which no reference
test exercises.

* An independent brute-force restatement of one group step of each solver
  (phaser/engines/conventional/solvers.py:171-290 and :415-504), written with explicit
  per-position / per-mode / per-pixel loops and a DFT matrix built from its definition
  (no numpy.fft, no oracle code), must agree with the oracle's FP64 twin to rounding:
  object, probe, per-position errors and position steps.
* Fixed point: when the data equal the model intensities (eps = 0), chi = 0 and a group
  step leaves object and probe unchanged; the dry-run rescale of a pure-phase object is 1
  (Parseval).
* S = 1 reduction: two slices whose second slice is transparent with a unit propagator
  give the single-slice step on slice 0 and the probe.
* The FP32 mirror agrees with the FP64 twin to FP32 rounding on a small case.
"""
import numpy
import pytest

import oracle.noise_models as onoise
import oracle.num as onum
import oracle.object as oobj
import oracle.state as ost
from oracle.solvers import Sim, epie_run, lsqml_dry_run, lsqml_run


# ---------------------------------------------------------------- brute force (FP64)

def _dft_mat(n):
    """fft2 convention of phaser (num.py:434-468): X = F a F^T with
    F[k, j] = exp(-2 pi i k (j - n/2) / n) / sqrt(n) (ifftshift folded in, even n)."""
    F = numpy.empty((n, n), complex)
    for k in range(n):
        for j in range(n):
            F[k, j] = numpy.exp(-2j * numpy.pi * k * (j - n // 2) / n) / numpy.sqrt(n)
    return F


class BF:
    def __init__(self, n):
        self.F = _dft_mat(n)
        self.G = self.F.conj().T   # inverse (fftshift folded in)

    def fft2(self, a):
        return self.F @ a @ self.F.T

    def ifft2(self, a):
        return self.G @ a @ self.G.T


def _freqs(n, dx):
    return numpy.array([(i if i < n / 2 else i - n) / (n * dx) for i in range(n)])


def _geometry(pos, corner, dx, n):
    """start (rint, half-even) and subpixel shift of every position."""
    starts, shifts = [], []
    for p in pos:
        idx = [(p[a] - corner[a]) / dx - (n - 1) / 2 for a in range(2)]
        st = [int(numpy.rint(v)) for v in idx]
        starts.append(st)
        shifts.append([(idx[a] - st[a]) * dx for a in range(2)])
    return starts, shifts


def _forward(bf, obj, probe, pos, corner, dx, props):
    S, H, W = obj.shape
    M, n, _ = probe.shape
    B = len(pos)
    k = _freqs(n, dx)
    starts, shifts = _geometry(pos, corner, dx, n)
    cut = numpy.empty((B, S, n, n), complex)
    for b in range(B):
        for s in range(S):
            for y in range(n):
                for x in range(n):
                    cut[b, s, y, x] = obj[s, starts[b][0] + y, starts[b][1] + x]
    filt = numpy.empty((B, n, n), complex)
    for b in range(B):
        for iy in range(n):
            for ix in range(n):
                filt[b, iy, ix] = numpy.exp(-2j * numpy.pi * (shifts[b][0] * k[iy] + shifts[b][1] * k[ix]))
    psi = numpy.empty((S, B, M, n, n), complex)
    for b in range(B):
        for m in range(M):
            psi[0, b, m] = bf.ifft2(bf.fft2(probe[m]) * filt[b])
    gpm = numpy.zeros((S, H, W))
    for s in range(S):
        for b in range(B):
            for m in range(M):
                for y in range(n):
                    for x in range(n):
                        gpm[s, starts[b][0] + y, starts[b][1] + x] += abs(psi[s, b, m, y, x])**2
                if s < S - 1:
                    psi[s + 1, b, m] = bf.ifft2(bf.fft2(psi[s, b, m] * cut[b, s]) * props[s])
    Psi = numpy.empty((B, M, n, n), complex)
    for b in range(B):
        for m in range(M):
            Psi[b, m] = bf.fft2(psi[S - 1, b, m] * cut[b, S - 1])
    inten = numpy.sum(numpy.abs(Psi)**2, axis=1)
    return dict(k=k, starts=starts, cut=cut, filt=filt, psi=psi, gpm=gpm, Psi=Psi, inten=inten)


def _chi(bf, f, D, mask, eps):
    B, M, n, _ = f['Psi'].shape
    chi = numpy.empty((B, M, n, n), complex)
    for b in range(B):
        u = (numpy.sqrt(numpy.maximum(D[b], 0.0)) / (numpy.sqrt(f['inten'][b]) + eps) - 1.0) * mask
        for m in range(M):
            chi[b, m] = bf.ifft2(u * f['Psi'][b, m])
    errors = numpy.array([numpy.sqrt(numpy.sum((f['inten'][b] - D[b])**2)) for b in range(B)])
    return chi, errors


def _pos_steps(bf, f, chi, dx):
    """solvers.py:275-288: per-position numerator, global denominator."""
    S, B, M, n, _ = f['psi'].shape
    k = f['k']
    ks = (numpy.repeat(k[:, None], n, axis=1), numpy.repeat(k[None, :], n, axis=0))
    out = numpy.zeros((B, 2))
    for a in range(2):
        num = numpy.zeros(B)
        den = 0.0
        for b in range(B):
            norm = sum(numpy.sum(numpy.abs(f['psi'][0, b, m])**2) for m in range(M))
            for m in range(M):
                dPx = bf.ifft2(bf.fft2(f['psi'][0, b, m]) / norm * (-2j * numpy.pi * ks[a]))
                prod = dPx * f['cut'][b, 0]
                num[b] += numpy.sum(numpy.real(chi[b, m] * numpy.conj(prod)))
                den += numpy.sum(numpy.abs(prod)**2)
        out[:, a] = num / den
    return out


def bf_lsqml(obj, probe, pos, corner, dx, props, D, mask, *, beta_o, beta_p, reg_o, reg_p, gamma, eps):
    obj = obj.copy()
    probe = probe.copy()
    S, H, W = obj.shape
    M, n, _ = probe.shape
    B = len(pos)
    bf = BF(n)
    f = _forward(bf, obj, probe, pos, corner, dx, props)
    chi, errors = _chi(bf, f, D, mask, eps)
    cut, psi, gpm, starts = f['cut'], f['psi'], f['gpm'], f['starts']
    gom = numpy.zeros((n, n))
    for b in range(B):
        prod = numpy.ones((n, n), complex)
        for s in range(S):
            prod = prod * cut[b, s]
        gom += numpy.abs(prod)**2
    for s in range(S - 1, -1, -1):
        dP = numpy.empty_like(chi)
        for b in range(B):
            for m in range(M):
                dP[b, m] = chi[b, m] * numpy.conj(cut[b, s])
        # object step: alpha_O,b = sum_{m,px} Re(chi conj(dO psi)) / (sum_all |dO psi|^2 + gamma)
        num = numpy.zeros(B)
        den = gamma
        acc = numpy.zeros((H, W), complex)
        for b in range(B):
            for m in range(M):
                dO = chi[b, m] * numpy.conj(psi[s, b, m])
                t = dO * psi[s, b, m]
                num[b] += numpy.sum(numpy.real(chi[b, m] * numpy.conj(t)))
                den += numpy.sum(numpy.abs(t)**2)
                for y in range(n):
                    for x in range(n):
                        acc[starts[b][0] + y, starts[b][1] + x] += dO[y, x]
        A = numpy.sum(num / den)
        g = gpm[s]
        obj[s] += beta_o * A * (acc / (g + reg_o)) * g / (g + 1e-16)
        if s > 0:
            for b in range(B):
                for m in range(M):
                    chi[b, m] = bf.ifft2(bf.fft2(dP[b, m]) * numpy.conj(props[s - 1]))
        else:
            denP = gamma
            numP = numpy.zeros((B, M))
            for b in range(B):
                for m in range(M):
                    t = dP[b, m] * cut[b, 0]
                    numP[b, m] = numpy.sum(numpy.real(chi[b, m] * numpy.conj(t)))
                    denP += numpy.sum(numpy.abs(t)**2)
            for m in range(M):
                spec = numpy.zeros((n, n), complex)
                for b in range(B):
                    spec += bf.fft2(dP[b, m]) * numpy.conj(f['filt'][b])
                dPavg = bf.ifft2(spec) / (gom + reg_p)
                probe[m] += beta_p * numpy.sum(numP[:, m] / denP) * dPavg * gom / (gom + 1e-16)
    return obj, probe, errors, _pos_steps(bf, f, chi, dx)


def bf_epie(obj, probe, pos, corner, dx, props, D, mask, *, beta_o, beta_p, eps):
    obj = obj.copy()
    probe = probe.copy()
    S, H, W = obj.shape
    M, n, _ = probe.shape
    B = len(pos)
    bf = BF(n)
    f = _forward(bf, obj, probe, pos, corner, dx, props)
    chi, errors = _chi(bf, f, D, mask, eps)
    cut, psi, starts = f['cut'], f['psi'], f['starts']
    for s in range(S - 1, -1, -1):
        pu = numpy.empty_like(chi)
        for b in range(B):
            mo = numpy.max(numpy.abs(cut[b, s])**2)
            mp = numpy.max(sum(numpy.abs(psi[s, b, m])**2 for m in range(M)))
            upd = numpy.zeros((n, n), complex)
            for m in range(M):
                pu[b, m] = numpy.conj(cut[b, s]) * chi[b, m] / mo
                upd += numpy.conj(psi[s, b, m]) * chi[b, m]
            upd *= beta_o / S / mp
            for y in range(n):
                for x in range(n):
                    obj[s, starts[b][0] + y, starts[b][1] + x] += upd[y, x]
        if s > 0:
            for b in range(B):
                for m in range(M):
                    chi[b, m] = bf.ifft2(bf.fft2(pu[b, m]) * numpy.conj(props[s - 1]))
        else:
            for m in range(M):
                spec = numpy.zeros((n, n), complex)
                for b in range(B):
                    spec += bf.fft2(pu[b, m]) * numpy.conj(f['filt'][b])
                probe[m] += beta_p * bf.ifft2(spec / B)
    return obj, probe, errors, _pos_steps(bf, f, chi, dx)


# ---------------------------------------------------------------- shared tiny problem

N, DX = 8, 0.25


def _tiny(S=2, M=2, B=3, seed=0):
    rng = numpy.random.default_rng(seed)
    H = W = 18
    obj = numpy.exp(1j * rng.uniform(-0.5, 0.5, (S, H, W))) * rng.uniform(0.8, 1.1, (S, H, W))
    probe = (rng.normal(size=(M, N, N)) + 1j * rng.normal(size=(M, N, N))) * 0.3
    corner = (0.0, 0.0)
    starts = rng.integers(1, H - N - 1, size=(B, 2))
    frac = rng.uniform(-0.4, 0.4, size=(B, 2))
    pos = (starts + (N - 1) / 2 + frac) * DX
    props = numpy.exp(1j * rng.uniform(-numpy.pi, numpy.pi, (max(S - 1, 0), N, N)))
    D = rng.uniform(0.0, 2.0, (B, N, N)) * numpy.sum(numpy.abs(probe)**2)
    mask = (rng.uniform(size=(N, N)) > 0.1).astype(float)
    return dict(obj=obj, probe=probe, corner=corner, pos=pos, props=props, D=D, mask=mask)


def _oracle_state(p, dtype=numpy.float64):
    cdt = numpy.complex128 if dtype == numpy.float64 else numpy.complex64
    S, H, W = p['obj'].shape
    probe = ost.ProbeState(onum.Sampling((N, N), sampling=(DX, DX)), p['probe'].astype(cdt))
    osamp = oobj.ObjectSampling((H, W), DX, corner=p['corner'])
    obj = ost.ObjectState(osamp, p['obj'].astype(cdt), numpy.full(S, 20.0, dtype=dtype))
    return ost.ReconsState(iter=ost.IterState(), wavelength=0.02, probe=probe, object=obj,
                           scan=numpy.asarray(p['pos'], dtype=dtype))


def _run(p, solver, dtype=numpy.float64, eps=1e-3, props=None, **kw):
    st = _oracle_state(p, dtype)
    sim = Sim(st, onoise.AmplitudeNoiseModel(eps=eps), dtype)
    B = len(p['pos'])
    group = numpy.arange(B)[None, :]
    props = p['props'] if props is None else props
    props = props.astype(numpy.complex128 if dtype == numpy.float64 else numpy.complex64) if len(props) else None
    pats = p['D'].astype(dtype)
    mask = p['mask'].astype(dtype)
    if solver == 'lsqml':
        err, pu = lsqml_run(sim, group, pats, pattern_mask=mask, props=props, update_position=True,
                            calc_error=True, **kw)
    else:
        err, pu = epie_run(sim, group, pats, pattern_mask=mask, props=props, update_position=True,
                           calc_error=True, **kw)
    return st.object.data, st.probe.data, err, pu


def _rel(a, b):
    return numpy.linalg.norm(numpy.ravel(a - b)) / numpy.linalg.norm(numpy.ravel(b))


LSQ = dict(beta_object=0.9, beta_probe=0.8, illum_reg_object=1e-2, illum_reg_probe=1e-2, gamma=1e-4)


@pytest.mark.parametrize('S,M', [(1, 1), (2, 2), (3, 1)])
def test_lsqml_step_matches_brute_force(S, M):
    p = _tiny(S=S, M=M, seed=S * 10 + M)
    o, pr, err, pu = _run(p, 'lsqml', **LSQ)
    bo, bp, berr, bpu = bf_lsqml(p['obj'], p['probe'], p['pos'], p['corner'], DX, p['props'], p['D'], p['mask'],
                                 beta_o=0.9, beta_p=0.8, reg_o=1e-2, reg_p=1e-2, gamma=1e-4, eps=1e-3)
    assert _rel(o - p['obj'], bo - p['obj']) < 1e-9    # the update itself, not just the sum
    assert _rel(pr - p['probe'], bp - p['probe']) < 1e-9
    assert _rel(err, berr) < 1e-12
    assert _rel(pu, bpu) < 1e-9


@pytest.mark.parametrize('S,M', [(1, 1), (2, 2)])
def test_epie_step_matches_brute_force(S, M):
    p = _tiny(S=S, M=M, seed=S * 7 + M)
    o, pr, err, pu = _run(p, 'epie', beta_object=0.7, beta_probe=0.6)
    bo, bp, berr, bpu = bf_epie(p['obj'], p['probe'], p['pos'], p['corner'], DX, p['props'], p['D'], p['mask'],
                                beta_o=0.7, beta_p=0.6, eps=1e-3)
    assert _rel(o - p['obj'], bo - p['obj']) < 1e-9
    assert _rel(pr - p['probe'], bp - p['probe']) < 1e-9
    assert _rel(err, berr) < 1e-12
    assert _rel(pu, bpu) < 1e-9


def test_brute_force_detects_a_planted_mistake():
    """The comparison is sensitive: a conjugate dropped in the oracle's probe step fails it."""
    import oracle.solvers as osol
    p = _tiny(S=2, M=2, seed=3)
    bo, bp, _, _ = bf_lsqml(p['obj'], p['probe'], p['pos'], p['corner'], DX, p['props'], p['D'], p['mask'],
                            beta_o=0.9, beta_p=0.8, reg_o=1e-2, reg_p=1e-2, gamma=1e-4, eps=1e-3)
    orig = osol.fourier_shift_filter
    try:
        osol.fourier_shift_filter = lambda ky, kx, sh: orig(ky, kx, sh).conj()
        _, pr, _, _ = _run(p, 'lsqml', **LSQ)
    finally:
        osol.fourier_shift_filter = orig
    assert _rel(pr - p['probe'], bp - p['probe']) > 1e-3


def _consistent_data(p):
    """Patterns equal to the model intensities of the state itself (FP64)."""
    bf = BF(N)
    f = _forward(bf, p['obj'], p['probe'], p['pos'], p['corner'], DX, p['props'])
    q = dict(p)
    q['D'] = f['inten']
    q['mask'] = numpy.ones((N, N))
    return q


@pytest.mark.parametrize('solver', ['lsqml', 'epie'])
def test_data_equal_model_is_a_fixed_point(solver):
    p = _consistent_data(_tiny(S=2, M=2, seed=11))
    kw = LSQ if solver == 'lsqml' else dict(beta_object=1.0, beta_probe=1.0)
    o, pr, err, pu = _run(p, solver, eps=0.0, **kw)
    assert numpy.max(numpy.abs(o - p['obj'])) < 1e-12
    assert numpy.max(numpy.abs(pr - p['probe'])) < 1e-12
    assert numpy.max(err) < 1e-9 * numpy.max(p['D'])
    assert numpy.max(numpy.abs(pu)) < 1e-9


def test_dry_run_rescale_is_one_for_pure_phase_object():
    p = _consistent_data(_tiny(S=2, M=2, seed=12))
    p['obj'] = numpy.exp(1j * numpy.angle(p['obj']))
    p = _consistent_data(p)
    st = _oracle_state(p)
    sim = Sim(st, onoise.AmplitudeNoiseModel(), numpy.float64)
    r = lsqml_dry_run(sim, numpy.arange(3)[None, :], p['D'], pattern_mask=p['mask'], props=p['props'])
    numpy.testing.assert_allclose(r, 1.0, atol=1e-12)


def test_transparent_second_slice_reduces_to_single_slice():
    p2 = _tiny(S=2, M=2, seed=13)
    p2['obj'][1] = 1.0
    p2['props'] = numpy.ones((1, N, N), complex)
    p1 = dict(p2)
    p1['obj'] = p2['obj'][:1].copy()
    p1['props'] = numpy.ones((0, N, N), complex)
    o2, pr2, e2, u2 = _run(p2, 'lsqml', **LSQ)
    o1, pr1, e1, u1 = _run(p1, 'lsqml', **LSQ)
    assert _rel(o2[0] - p2['obj'][0], o1[0] - p1['obj'][0]) < 1e-10
    assert _rel(pr2 - p2['probe'], pr1 - p1['probe']) < 1e-10
    assert _rel(e2, e1) < 1e-12
    assert _rel(u2, u1) < 1e-10


def test_fp32_mirror_agrees_with_fp64_twin():
    p = _tiny(S=2, M=2, seed=14)
    o64, p64, e64, u64 = _run(p, 'lsqml', numpy.float64, **LSQ)
    o32, p32, e32, u32 = _run(p, 'lsqml', numpy.float32, **LSQ)
    assert o32.dtype == numpy.complex64 and p32.dtype == numpy.complex64
    assert _rel(o32, o64) < 1e-5
    assert _rel(p32, p64) < 1e-5
    assert _rel(e32, e64) < 1e-5
