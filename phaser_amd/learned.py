"""Learned initializer: an MLP that maps a binned diffraction pattern to the phase and
amplitude of the object patch under the beam, trained on reconstructions the engines
themselves produce, whose per-patch guesses are stitched into the starting object.

There is no counterpart in the reference.  Everything on the GPU runs through this
package's kernels (phx_mlp.hip): BF16 MFMA GEMMs with FP32 accumulation for the forward
and backward passes, Adam on FP32 master weights with BF16 copies, the pattern features,
the patch targets and the stitching.  torch only allocates.

Network: input log(1 + 2x2-binned pattern) (n^2/4 values; 4096 for 128^2 patterns),
dense layers with ReLU between them, linear output [phase (ps^2), amplitude (ps^2)] of
the projected object prod_s O_s over the ps x ps window centred on the beam.  Loss: mean
squared error.  The reconstruction engines decide the result: `nrmse` measures any start
or result against a reference object, and a caller falls back to the flat start when
the network does not improve on it.
"""
import ctypes
import math
import time

import numpy

from . import _lib

DEFAULT_SIZES = (4096, 2048, 2048, 2048, 2048)
PATCH = 32


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned initializer runs on the GPU (no CPU fallback)")
    return torch


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _pad128(n):
    return (n + 127) // 128 * 128


def gemm_nt(A, B, M, N, K, *, C=None, Cb=None, Ct=None, bias=None, relu=False, mask=None, accumulate=False):
    """C/Cb[M, N] = A[M, K] B[N, K]^T (+ epilogue), Ct = its BF16 transpose, through
    phx_gemm_bf16_nt (both operands k-contiguous)."""
    ldc = 0 if C is None and Cb is None else (C if C is not None else Cb).shape[-1]
    _lib.call("phx_gemm_bf16_nt", _p(A), A.shape[-1], _p(B), B.shape[-1], M, N, K, _p(C), _p(Cb), ldc, _p(Ct),
              0 if Ct is None else Ct.shape[-1], _p(bias), int(relu), _p(mask), 0 if mask is None else mask.shape[-1],
              int(accumulate), _stream())


def transpose_bf16(A, out):
    """out (cols x rows) = A^T for a rows x cols BF16 tensor."""
    _lib.call("phx_transpose_bf16", _p(A), A.shape[-1], A.shape[0], A.shape[-1], _p(out), out.shape[-1], _stream())


def gemm(A, B, M, N, K, *, trans_a=False, trans_b=False, C=None, Cb=None, bias=None, relu=False, mask=None,
         accumulate=False):
    """C/Cb[M, N] = op(A) op(B) (+ epilogue), through phx_gemm_bf16."""
    lda = A.shape[-1]
    ldb = B.shape[-1]
    ldc = (C if C is not None else Cb).shape[-1]
    ldm = 0 if mask is None else mask.shape[-1]
    _lib.call("phx_gemm_bf16", _p(A), lda, int(trans_a), _p(B), ldb, int(trans_b), M, N, K, _p(C), _p(Cb), ldc,
              _p(bias), int(relu), _p(mask), ldm, int(accumulate), _stream())


class MLP:
    """Dense ReLU network with BF16 weights (FP32 masters), trained with Adam."""

    def __init__(self, sizes=DEFAULT_SIZES, seed=0, device='cuda'):
        torch = _torch()
        _lib.load()
        for s in sizes:
            if s % 128:
                raise ValueError(f"layer sizes must be multiples of 128 (got {sizes})")
        self.sizes = tuple(int(s) for s in sizes)
        self.L = len(sizes) - 1
        self.device = device
        rng = numpy.random.default_rng(seed)
        # FP32 master weights W (in x out), their BF16 copy Wb and its transpose Wbt
        # (out x in): every GEMM takes the k-contiguous NT form
        self.W, self.Wb, self.Wbt, self.b = [], [], [], []
        for l, (a, o) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            # He init; the linear output layer starts 10x smaller, or the first steps drive
            # the last hidden layer's ReLUs dead (its active fraction falls to ~5%)
            scale = math.sqrt(2.0 / a) * (0.1 if l == len(sizes) - 2 else 1.0)
            w = (rng.standard_normal((a, o)) * scale).astype(numpy.float32)
            self.W.append(torch.from_numpy(w).to(device))
            self.Wb.append(torch.empty((a, o), dtype=torch.bfloat16, device=device))
            self.Wbt.append(torch.empty((o, a), dtype=torch.bfloat16, device=device))
            _lib.call("phx_f32_to_bf16", _p(self.W[-1]), _p(self.Wb[-1]), a * o, _stream())
            transpose_bf16(self.Wb[-1], self.Wbt[-1])
            self.b.append(torch.zeros(o, dtype=torch.float32, device=device))
        self.dW = [torch.zeros_like(w) for w in self.W]
        self.db = [torch.zeros_like(b) for b in self.b]
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.t = 0
        self._bufs = {}
        self.loss_acc = torch.zeros(1, dtype=torch.float64, device=device)

    def _buffers(self, batch):
        if batch not in self._bufs:
            torch = _torch()
            bf, f32 = torch.bfloat16, torch.float32
            acts = [None] + [torch.empty((batch, s), dtype=bf, device=self.device) for s in self.sizes[1:-1]]
            # transposed activations (size x batch) and output gradients: the weight
            # gradients' k-contiguous operands
            actsT = [torch.empty((s, batch), dtype=bf, device=self.device) for s in self.sizes[:-1]]
            out = torch.empty((batch, self.sizes[-1]), dtype=f32, device=self.device)
            dz = [torch.empty((batch, s), dtype=f32, device=self.device) for s in self.sizes[1:]]
            dzb = [torch.empty((batch, s), dtype=bf, device=self.device) for s in self.sizes[1:]]
            dzbT = [torch.empty((s, batch), dtype=bf, device=self.device) for s in self.sizes[1:]]
            self._bufs[batch] = (acts, actsT, out, dz, dzb, dzbT)
        return self._bufs[batch]

    def layer_forward(self, l, A, C=None, Cb=None, Ct=None, relu=True):
        """Z = A W_l + b_l (ReLU'd with relu) into C (FP32), Cb (BF16) and/or its BF16
        transpose Ct."""
        gemm_nt(A, self.Wbt[l], A.shape[0], self.sizes[l + 1], self.sizes[l], C=C, Cb=Cb, Ct=Ct, bias=self.b[l],
                relu=relu)

    def forward(self, X, train=True):
        """X: (batch, sizes[0]) BF16, batch a multiple of 128.  Returns the FP32 output
        buffer (reused by the next call with the same batch).  train=True also keeps the
        transposed activations backward() needs."""
        batch = X.shape[0]
        acts, actsT, out, _, _, _ = self._buffers(batch)
        acts[0] = X
        if train:
            transpose_bf16(X, actsT[0])
        for l in range(self.L):
            last = l == self.L - 1
            self.layer_forward(l, acts[l], C=out if last else None, Cb=None if last else acts[l + 1],
                               Ct=actsT[l + 1] if train and not last else None, relu=not last)
        return out

    def backward(self, T):
        """Gradients of mean((out - T)^2) after forward(); the loss sum accumulates in
        self.loss_acc (FP64)."""
        batch = T.shape[0]
        acts, actsT, out, dz, dzb, dzbT = self._buffers(batch)
        n = out.numel()
        _lib.call("phx_mse_grad_t", _p(out), _p(T), batch, self.sizes[-1], 2.0 / n, _p(dz[-1]), _p(dzb[-1]),
                  _p(dzbT[-1]), _p(self.loss_acc), _stream())
        for l in range(self.L - 1, -1, -1):
            # dW_l = A_l^T dZ_l: (in x batch) times (out x batch)^T
            gemm_nt(actsT[l], dzbT[l], self.sizes[l], self.sizes[l + 1], batch, C=self.dW[l])
            _lib.call("phx_colsum_f32", _p(dz[l]), batch, self.sizes[l + 1], self.sizes[l + 1], _p(self.db[l]),
                      _stream())
            if l > 0:
                # dZ_{l-1} = (dZ_l W_l^T) * relu'(Z_{l-1}): W_l (in x out) is the N x K operand
                gemm_nt(dzb[l], self.Wb[l], batch, self.sizes[l], self.sizes[l + 1], C=dz[l - 1], Cb=dzb[l - 1],
                        Ct=dzbT[l - 1], mask=acts[l])

    def step(self, lr, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
        self.t += 1
        c1 = 1.0 / (1.0 - b1**self.t)
        c2 = 1.0 / (1.0 - b2**self.t)
        st = _stream()
        for l in range(self.L):
            _lib.call("phx_adam_f32_bf16", _p(self.W[l]), _p(self.Wb[l]), _p(self.dW[l]), _p(self.mW[l]),
                      _p(self.vW[l]), self.W[l].numel(), lr, b1, b2, eps, c1, c2, weight_decay, st)
            transpose_bf16(self.Wb[l], self.Wbt[l])
            _lib.call("phx_adam_f32_bf16", _p(self.b[l]), None, _p(self.db[l]), _p(self.mb[l]), _p(self.vb[l]),
                      self.b[l].numel(), lr, b1, b2, eps, c1, c2, 0.0, st)

    def predict(self, X, batch=1024):
        """Outputs for every row of X (BF16, any count), FP32 (rows, out)."""
        torch = _torch()
        rows = X.shape[0]
        res = torch.empty((rows, self.sizes[-1]), dtype=torch.float32, device=self.device)
        for r0 in range(0, rows, batch):
            r1 = min(rows, r0 + batch)
            nb = _pad128(r1 - r0)
            if nb != r1 - r0:
                xb = torch.zeros((nb, X.shape[1]), dtype=X.dtype, device=self.device)
                xb[:r1 - r0] = X[r0:r1]
            else:
                xb = X[r0:r1]
            res[r0:r1] = self.forward(xb, train=False)[:r1 - r0]
        return res

    def state_dict(self):
        return {'sizes': self.sizes, 'W': [w.cpu() for w in self.W], 'b': [b.cpu() for b in self.b]}

    def load_state_dict(self, d):
        for l in range(self.L):
            self.W[l].copy_(d['W'][l])
            _lib.call("phx_f32_to_bf16", _p(self.W[l]), _p(self.Wb[l]), self.W[l].numel(), _stream())
            transpose_bf16(self.Wb[l], self.Wbt[l])
            self.b[l].copy_(d['b'][l])


# ----------------------------------------------------------------------------- data

def feature_bin(n):
    """Pattern binning that gives the network's 64 x 64 input (2 at 128^2, 3 at 192^2)."""
    if n % 64:
        raise ValueError(f"pattern side {n} is not a multiple of 64")
    return n // 64


def features(patterns_dev, n, index=None, inv_scale=1.0, bin=2):
    """(count, (n/bin)^2) BF16 network inputs from device patterns (P, n, n) FP32."""
    torch = _torch()
    count = patterns_dev.shape[0] if index is None else index.shape[0]
    out = torch.empty((count, (n // bin)**2), dtype=torch.bfloat16, device=patterns_dev.device)
    _lib.call("phx_pattern_features", _p(patterns_dev), _p(index), count, n, bin, float(inv_scale), _p(out),
              _stream())
    return out


def window_starts(obj_sampling, scan, n, ps=PATCH):
    """(P, 2) int32 start of the ps x ps window centred on each beam position."""
    start, _shift = obj_sampling.cutout_indices(numpy.asarray(scan).reshape(-1, 2), (n, n))
    return (start + (n - ps) // 2).astype(numpy.int32)


def normalize_object(obj):
    """Fix the reconstruction's global phase and amplitude (both free in ptychography):
    slice 0 is scaled so the projected object has mean phase 0 and mean amplitude 1
    (over all pixels passed in)."""
    proj = numpy.prod(obj.astype(numpy.complex128), axis=0)
    ph = numpy.angle(numpy.sum(proj))
    amp = numpy.mean(numpy.abs(proj))
    out = obj.copy()
    out[0] = out[0] * numpy.exp(-1j * ph) / amp
    return out


def patch_targets(obj_dev, starts_dev, ps=PATCH):
    """(P, 2 ps^2) FP32 [phase, amplitude] of the projected object over each window."""
    torch = _torch()
    S, H, W = (int(v) for v in obj_dev.shape)
    count = starts_dev.shape[0]
    out = torch.empty((count, 2 * ps * ps), dtype=torch.float32, device=obj_dev.device)
    _lib.call("phx_object_patches", _p(obj_dev), S, H, W, _p(starts_dev), count, ps, 0.0, _p(out), _stream())
    return out


def stitch(pred, starts_dev, S, H, W, ps=PATCH):
    """Average the predicted windows into an (S, H, W) complex64 object (phase and
    log-amplitude shared equally by the slices; 1 where no window lands)."""
    torch = _torch()
    dev = pred.device
    acc = torch.zeros((3, H * W), dtype=torch.float32, device=dev)
    obj = torch.empty((S, H, W), dtype=torch.complex64, device=dev)
    _lib.call("phx_stitch_patches", _p(pred), pred.shape[1], _p(starts_dev), starts_dev.shape[0], ps, S, H, W,
              _p(acc[0]), _p(acc[1]), _p(acc[2]), _p(obj), _stream())
    return obj


def nrmse(obj, truth, crop=None):
    """Object NRMSE with the optimal global phase over the crop, on the projected object
    prod_s O_s: min_phi ||a e^{i phi} - b|| / ||b||."""
    a = numpy.prod(numpy.asarray(obj, dtype=numpy.complex128), axis=0)
    b = numpy.prod(numpy.asarray(truth, dtype=numpy.complex128), axis=0)
    if crop is not None:
        a, b = a[crop], b[crop]
    phi = numpy.angle(numpy.vdot(a, b))
    return float(numpy.linalg.norm(a * numpy.exp(1j * phi) - b) / numpy.linalg.norm(b))


def nrmse_tensor(obj, truth, crop=None):
    """nrmse() on torch tensors where they live (the GPU for a resident object), in
    complex128: no host copy of a large multislice object."""
    import torch
    sel = (slice(None),) + (tuple(crop) if crop is not None else ())
    a, b = obj[sel], truth[sel]
    pa = a[0].to(torch.complex128)
    pb = b[0].to(torch.complex128)
    for s in range(1, a.shape[0]):
        pa = pa * a[s]
        pb = pb * b[s]
    phi = torch.angle(torch.sum(torch.conj(pa) * pb))
    return float(torch.linalg.vector_norm(pa * torch.exp(1j * phi) - pb) / torch.linalg.vector_norm(pb))


def reconstruct_plain(state, pats, *, niter, grouping=64, seed=0):
    """Plain LSQ-ML (the engines' defaults) from the given state; returns the result."""
    from .engines.conventional import run_engine
    from .plan import ConventionalEnginePlan
    plan = ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=niter, grouping=grouping)
    args = {'data': pats, 'state': state, 'dtype': numpy.float32, 'seed': seed, 'observer': None, 'recons_name': ''}
    return run_engine(args, plan)


def reconstruct_gradient(state, pats, *, niter, grouping=64, seed=0, noise_model='poisson', lr_object=20.0,
                         momentum=0.5, group_constraints=(), observer=None):
    """The gradient engine (hand-written adjoint) with momentum SGD on the object only,
    from the given state; returns the result."""
    from .engines import gradient as geng
    from .plan import GradientEnginePlan
    plan = GradientEnginePlan(noise_model=noise_model, niter=niter, grouping=grouping,
                              solvers={'object': {'type': 'sgd', 'learning_rate': lr_object, 'momentum': momentum}},
                              group_constraints=list(group_constraints))
    args = {'data': pats, 'state': state.to_device(), 'dtype': numpy.float32, 'seed': seed, 'observer': observer,
            'recons_name': ''}
    return geng.run_engine(args, plan)


def training_examples(cfg, *, niter, method='lsqml', gradient_opts=None, log=None):
    """One synthetic object reconstructed by the method (plain LSQ-ML, or with
    method='gradient' the gradient engine with `gradient_opts`): (features BF16, targets
    FP32) on the device, targets taken from the normalized reconstruction."""
    torch = _torch()
    from .synthetic import make_dataset
    state, pats, _truth = make_dataset(cfg)
    if method == 'lsqml':
        out = reconstruct_plain(state, pats, niter=niter, seed=cfg.seed).to_numpy()
    elif method == 'gradient':
        out = reconstruct_gradient(state, pats, niter=niter, seed=cfg.seed, **(gradient_opts or {})).to_numpy()
    else:
        raise ValueError(f"unknown reconstruction method '{method}'")
    obj = normalize_object(out.object.data)
    n = cfg.n
    dev_p = torch.from_numpy(numpy.ascontiguousarray(pats.patterns.reshape(-1, n, n))).cuda()
    X = features(dev_p, n, bin=feature_bin(n))
    starts = torch.from_numpy(window_starts(out.object.sampling, out.scan, n)).cuda()
    T = patch_targets(torch.from_numpy(obj.astype(numpy.complex64)).cuda(), starts)
    return X, T


def train(mlp, X, T, *, epochs, batch=1024, lr=1e-4, seed=0, log=None, standardize=True):
    """Adam over shuffled minibatches (rows beyond the last full batch are dropped per
    epoch); returns the mean loss of each epoch.  With `standardize`, the network learns
    (T - mean) / std per output column and the affine map is folded into its last layer
    afterwards (W' = W diag(std), b' = b std + mean), so predictions are in target units."""
    torch = _torch()
    if standardize:
        mu = T.mean(dim=0)
        sd = T.std(dim=0).clamp_min(1e-6)
        hist = train(mlp, X, (T - mu) / sd, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log,
                     standardize=False)
        L = mlp.L - 1
        mlp.W[L].mul_(sd)
        mlp.b[L].mul_(sd).add_(mu)
        _lib.call("phx_f32_to_bf16", _p(mlp.W[L]), _p(mlp.Wb[L]), mlp.W[L].numel(), _stream())
        transpose_bf16(mlp.Wb[L], mlp.Wbt[L])
        return hist
    rows = X.shape[0]
    batch = min(batch, rows // 128 * 128)
    if batch < 128:
        raise ValueError("need at least 128 training rows")
    gen = torch.Generator(device='cpu').manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(rows, generator=gen).cuda()
        mlp.loss_acc.zero_()
        nb = rows // batch
        for i in range(nb):
            idx = perm[i * batch:(i + 1) * batch]
            mlp.forward(X[idx])
            mlp.backward(T[idx])
            mlp.step(lr)
        hist.append(float(mlp.loss_acc.item()) / (nb * batch * T.shape[1]))
        if log:
            log(f"epoch {ep + 1}: loss {hist[-1]:.5f}")
    return hist


def learned_object(mlp, pats, state, ps=PATCH, phase_only=False):
    """The stitched starting object (S, H, W) complex64 on the device for `state`'s scan
    and object grid, from the network's guesses on every pattern.  With `phase_only` the
    predicted amplitudes are replaced by 1 (a phase object: the guess carries only the
    network's patch phases)."""
    torch = _torch()
    n = int(state.probe.data.shape[-1])
    S, H, W = (int(v) for v in state.object.data.shape)
    if hasattr(pats.patterns, 'is_cuda') and pats.patterns.is_cuda:
        dev_p = pats.patterns.reshape(-1, n, n)
    else:
        dev_p = torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(pats.patterns).reshape(-1, n, n))).cuda()
    X = features(dev_p, n, bin=feature_bin(n))
    pred = mlp.predict(X)
    if phase_only:
        pred[:, ps * ps:] = 1.0
    starts = torch.from_numpy(window_starts(state.object.sampling, state.scan, n, ps)).cuda()
    return stitch(pred, starts, S, H, W, ps)


# ----------------------------------------------------------------------- experiment

def evaluation_config(seed=7):
    """The time-to-NRMSE problem: 128^2 patterns, 4 modes, 5 slices of 0.6 rad Gaussian
    atoms, 32 x 32 scan (step 6 px): flat-start NRMSE ~0.13, plain LSQ-ML needs ~28
    iterations to reach 0.05."""
    from .synthetic import SyntheticConfig
    return SyntheticConfig(n=128, n_modes=4, n_slices=5, scan_shape=(32, 32), scan_step_px=6.0, jitter_px=0.5,
                           seed=seed, atom_peak_rad=0.6)

class NrmseWatch:
    """Observer with the engines' callback names: after every iteration, the object's
    NRMSE against `truth` over `crop`; raises EarlyTermination at `threshold`.  It needs
    no per-group callbacks, so the engines run their captured-graph iterations."""
    wants_groups = False

    def __init__(self, truth, crop, threshold, t0):
        self.truth, self.crop, self.threshold, self.t0 = truth, crop, threshold, t0
        self.paused = 0.0   # time spent judging, kept off the clock
        self.history = []   # (seconds since t0 excluding judging, iteration, nrmse)

    def init_engine(self, *a, **k):
        pass

    def start_engine(self, *a, **k):
        pass

    def update_group(self, *a, **k):
        pass

    def finish_engine(self, *a, **k):
        pass

    def check(self, obj, it):
        torch = _torch()
        torch.cuda.synchronize()
        t = time.perf_counter()
        if hasattr(obj, 'is_cuda') and obj.is_cuda:
            if not hasattr(self, 'truth_dev'):
                self.truth_dev = torch.from_numpy(numpy.ascontiguousarray(self.truth)).to(obj.device)
            e = nrmse_tensor(obj.detach(), self.truth_dev, self.crop)
        else:
            e = nrmse(obj, self.truth, self.crop)
        self.history.append((t - self.t0 - self.paused, it, e))
        self.paused += time.perf_counter() - t
        return e

    def update_iteration(self, state, i, n, error=None):
        from .state import EarlyTermination
        if self.check(state.object.data, i) <= self.threshold:
            raise EarlyTermination(state)


def _watch_run(run, watch):
    """Run one arm under `watch`; an engine's non-finite-state error (the reference's
    ValueError naming the first bad group) ends the arm and is recorded as `diverged`."""
    from .state import EarlyTermination
    diverged = None
    try:
        run(watch)
    except EarlyTermination:
        pass
    except ValueError as e:
        if 'NaN or inf' not in str(e):
            raise
        diverged = str(e)
    hit = [h for h in watch.history if h[2] <= watch.threshold]
    return {'seconds': hit[0][0] if hit else None, 'iterations': hit[0][1] if hit else None,
            'final_nrmse': watch.history[-1][2] if watch.history else None, 'diverged': diverged,
            'history': [(round(t, 4), i, round(e, 5)) for t, i, e in watch.history]}


def time_to_nrmse(cfg, mlp, *, threshold=0.05, max_iter=60, grouping=64, lr_object=10.0, momentum=0.9, log=None,
                  data=None, gradient_opts=None, plain_plan=None, phase_only_init=False):
    """Wall time to object NRMSE <= threshold against the synthetic truth, on the problem
    `cfg` (patterns resident before the clock starts):
      plain:   LSQ-ML from the flat start;
      learned: the network's stitched guess (inference included in the time), then the
               gradient engine (hand-written adjoint; SGD with Nesterov momentum on the
               object: per-pixel normalised steps (Adam) amplify the weakly illuminated
               edge and stall near NRMSE 0.065 on these objects);
      flat:    the same gradient engine from the flat start (what the network adds).
    NRMSE: projected object prod_s O_s over the scanned region, optimal global phase.
    `gradient_opts` (reconstruct_gradient's keywords) replaces the gradient arms' default
    recipe (amplitude noise model, SGD lr_object / momentum); `plain_plan(niter)` gives the
    plain arm's ConventionalEnginePlan (default: LSQ-ML with the engine defaults);
    `phase_only_init` starts the learned arm from the network's phases alone (learned_object).  An arm whose
    state goes non-finite is recorded as diverged."""
    torch = _torch()
    from .synthetic import make_dataset
    state0, pats, truth = data if data is not None else make_dataset(cfg)
    # the patterns live in HBM before any clock starts (every run reads them there)
    from dataclasses import replace
    pats = replace(pats, patterns=torch.from_numpy(numpy.ascontiguousarray(pats.patterns, dtype=numpy.float32)).cuda())
    crop = state0.object.sampling.get_region_crop()
    res = {'threshold': threshold, 'workload': {
        'n': cfg.n, 'modes': cfg.n_modes, 'slices': cfg.n_slices, 'scan': list(cfg.scan_shape),
        'object': list(state0.object.data.shape), 'patterns': int(numpy.prod(cfg.scan_shape))}}
    res['flat_start_nrmse'] = nrmse(state0.object.data, truth, crop)

    gopts = gradient_opts if gradient_opts is not None else {
        'noise_model': 'amplitude', 'lr_object': lr_object, 'momentum': momentum}
    res['gradient_recipe'] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in gopts.items()}

    def gradient_run(st, watch, niter):
        reconstruct_gradient(st, pats, niter=niter, grouping=grouping, seed=0, observer=watch, **gopts)

    def make_plain_plan(niter):
        from .plan import ConventionalEnginePlan
        if plain_plan is not None:
            return plain_plan(niter)
        return ConventionalEnginePlan(noise_model='amplitude', solver='lsqml', niter=niter, grouping=grouping)

    def plain_run(st, watch, niter):
        from .engines.conventional import run_engine
        run_engine({'data': pats, 'state': st, 'dtype': numpy.float32, 'seed': 0, 'observer': watch,
                    'recons_name': ''}, make_plain_plan(niter))

    # warm every path once (kernel loading, first-use costs) outside the clock; a plain
    # recipe that diverges on this problem does so here too
    try:
        plain_run(state0.copy(), None, 1)
    except ValueError as e:
        if 'NaN or inf' not in str(e):
            raise
    gradient_run(state0.copy().to_device(), None, 1)
    learned_object(mlp, pats, state0, phase_only=phase_only_init)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    w = NrmseWatch(truth, crop, threshold, t0)
    res['plain_lsqml'] = _watch_run(lambda watch: plain_run(state0.copy(), watch, max_iter), w)
    if log:
        log(f"plain LSQ-ML: {res['plain_lsqml']['seconds']} s, final NRMSE {res['plain_lsqml']['final_nrmse']}"
            + (f", diverged: {res['plain_lsqml']['diverged']}" if res['plain_lsqml']['diverged'] else ""))

    t0 = time.perf_counter()
    w = NrmseWatch(truth, crop, threshold, t0)

    def learned(watch):
        st = state0.copy().to_device()
        st.object.data = learned_object(mlp, pats, st, phase_only=phase_only_init).contiguous()
        watch.check(st.object.data, 0)
        res['learned_init_nrmse'] = watch.history[-1][2]
        gradient_run(st, watch, max_iter)
    res['learned_gradient'] = _watch_run(learned, w)
    if log:
        log(f"learned + gradient: {res['learned_gradient']['seconds']} s, init NRMSE "
            f"{res.get('learned_init_nrmse')}, final {res['learned_gradient']['final_nrmse']}")

    # ablation: the same gradient engine from the flat start (what the network adds)
    t0 = time.perf_counter()
    w = NrmseWatch(truth, crop, threshold, t0)
    res['flat_gradient'] = _watch_run(lambda watch: gradient_run(state0.copy().to_device(), watch, max_iter), w)
    if log:
        log(f"flat + gradient: {res['flat_gradient']['seconds']} s, final {res['flat_gradient']['final_nrmse']}")
    a, b = res['plain_lsqml']['seconds'], res['learned_gradient']['seconds']
    res['ratio'] = (b / a) if (a and b) else None
    return res


def train_on_reconstructions(n_objects, *, base=None, niter=10, epochs=20, lr=1e-4, batch=1024, seed=0,
                             sizes=DEFAULT_SIZES, method='lsqml', gradient_opts=None, vary_slices=True, log=None):
    """Build the training set from reconstructions (plain LSQ-ML, or the gradient engine
    with method='gradient') of `n_objects` seeded synthetic objects (3-5 slices, or the
    base configuration's slice count without `vary_slices`), then train a fresh network.
    Returns (mlp, info)."""
    torch = _torch()
    from dataclasses import replace
    from .synthetic import SyntheticConfig
    base = base or SyntheticConfig(n=128, n_modes=4, scan_shape=(16, 16), scan_step_px=6.0, jitter_px=0.5,
                                   atom_peak_rad=0.6)
    t0 = time.perf_counter()
    Xs, Ts = [], []
    for k in range(n_objects):
        cfg = replace(base, n_slices=(3 + k % 3) if vary_slices else base.n_slices, seed=seed * 100003 + 1000 + k)
        X, T = training_examples(cfg, niter=niter, method=method, gradient_opts=gradient_opts)
        Xs.append(X)
        Ts.append(T)
        if log and (k + 1) % 100 == 0:
            log(f"training set: {k + 1} of {n_objects} objects reconstructed ({time.perf_counter() - t0:.0f} s)")
    X = torch.cat(Xs)
    T = torch.cat(Ts)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    if log:
        log(f"training set: {X.shape[0]} patterns from {n_objects} reconstructions in {t_data:.1f} s")
    mlp = MLP(sizes, seed=seed)
    t0 = time.perf_counter()
    hist = train(mlp, X, T, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    return mlp, {'rows': int(X.shape[0]), 'objects': n_objects, 'data_seconds': t_data, 'train_seconds': t_train,
                 'epochs': epochs, 'loss_first_last': [hist[0], hist[-1]], 'method': method,
                 'reconstruction_iterations': niter}
