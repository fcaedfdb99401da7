"""MLP of the learned initializer: forward, MSE loss, backward and Adam, restated with
numpy.  Test infrastructure only.

There is no reference counterpart (the initializer is new).  The restatement follows
the textbook definitions: dense layers Z_l = A_{l-1} W_l + b_l with ReLU between them and
a linear last layer, loss = mean((Z_L - T)^2), reverse-mode gradients, and Adam (Kingma &
Ba 2015, bias-corrected moments).

`bf16=True` mirrors where the HIP kernels round to bfloat16 (round to nearest even): the
network input, the weights, every stored activation and every gradient that feeds a
GEMM; sums are exact (FP64) here and FP32 on the GPU.  With `bf16=False` everything is
FP64, which is what the finite-difference and autograd checks pin.
"""
import numpy


def round_bf16(a):
    """float -> nearest bfloat16 (ties to even), returned as float32."""
    a = numpy.ascontiguousarray(a, dtype=numpy.float32)
    u = a.view(numpy.uint32).astype(numpy.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(numpy.uint32)
    out = r.view(numpy.float32).copy()
    nan = numpy.isnan(a)
    out[nan] = a[nan]
    return out


def _r(a, bf16):
    return round_bf16(a).astype(numpy.float64) if bf16 else numpy.asarray(a, dtype=numpy.float64)


def forward(X, Ws, bs, *, bf16=False):
    """Returns (output Z_L, list of stored activations A_0 .. A_{L-1})."""
    A = _r(X, bf16)
    acts = [A]
    L = len(Ws)
    for l in range(L):
        Z = A @ _r(Ws[l], bf16) + numpy.asarray(bs[l], dtype=numpy.float64)
        if l < L - 1:
            A = _r(numpy.maximum(Z, 0.0), bf16)
            acts.append(A)
        else:
            return Z, acts


def loss(Z, T):
    return float(numpy.mean((Z - T)**2))


def backward(Z, T, acts, Ws, *, bf16=False):
    """Gradients (dWs, dbs) of loss(Z, T) = mean((Z - T)^2)."""
    L = len(Ws)
    dZ = 2.0 * (Z - T) / Z.size
    dWs, dbs = [None] * L, [None] * L
    for l in range(L - 1, -1, -1):
        dZr = _r(dZ, bf16)
        dWs[l] = acts[l].T @ dZr
        dbs[l] = dZ.sum(axis=0)
        if l > 0:
            dA = dZr @ _r(Ws[l], bf16).T
            dZ = dA * (acts[l] > 0)
    return dWs, dbs


def adam(w, g, m, v, t, *, lr, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
    """One Adam step (in place on w, m, v); t is the step number starting at 1."""
    g = g + weight_decay * w
    m[...] = b1 * m + (1 - b1) * g
    v[...] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1**t)
    vh = v / (1 - b2**t)
    w[...] = w - lr * mh / (numpy.sqrt(vh) + eps)
    return w


def features(patterns, inv_scale, bin=2):
    """log(1 + (bin x bin)-binned pattern * inv_scale), flattened rows."""
    P = numpy.asarray(patterns, dtype=numpy.float64)
    n = P.shape[-1]
    b = P.reshape(*P.shape[:-2], n // bin, bin, n // bin, bin).sum(axis=(-3, -1))
    return numpy.log1p(numpy.maximum(b, 0.0) * inv_scale).reshape(*P.shape[:-2], -1)


def nrmse(obj, truth, crop=None):
    """min over a global phase phi of ||prod_s O_s e^{i phi} - prod_s T_s|| / ||prod_s T_s||
    over the crop (the optimal phi = angle(sum conj(a) b))."""
    a = numpy.prod(numpy.asarray(obj, dtype=numpy.complex128), axis=0)
    b = numpy.prod(numpy.asarray(truth, dtype=numpy.complex128), axis=0)
    if crop is not None:
        a, b = a[crop], b[crop]
    phi = numpy.angle(numpy.vdot(a, b))
    return float(numpy.linalg.norm(a * numpy.exp(1j * phi) - b) / numpy.linalg.norm(b))
