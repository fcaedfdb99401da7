"""Learned-initializer kernels on the GPU against the CPU oracle (oracle/mlp.py):
BF16 MFMA GEMMs in all four transpose forms with their epilogues, the double-buffered NT
kernel with its transposed output, the BF16 transpose, a full training step
(forward, backward, Adam) of a small MLP, the features, patch targets and stitching, and
training that lowers the loss."""
import numpy
import pytest

import oracle.mlp as om
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _bf(torch, a):
    return torch.from_numpy(numpy.ascontiguousarray(a, dtype=numpy.float32)).cuda().to(torch.bfloat16)


@pytest.mark.parametrize('ta,tb', [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_bf16_transposes(ta, tb):
    import torch
    from phaser_amd.learned import gemm
    M, N, K = 256, 384, 96
    rng = numpy.random.default_rng(int(ta) * 2 + int(tb))
    A = om.round_bf16(rng.normal(size=(K, M) if ta else (M, K)))
    B = om.round_bf16(rng.normal(size=(N, K) if tb else (K, N)))
    ref = (A.T if ta else A).astype(numpy.float64) @ (B.T if tb else B).astype(numpy.float64)
    C = torch.empty((M, N), dtype=torch.float32, device='cuda')
    gemm(_bf(torch, A), _bf(torch, B), M, N, K, trans_a=ta, trans_b=tb, C=C)
    err = rel_err(C.cpu().numpy(), ref)
    print(f"gemm ta={ta} tb={tb}: rel err {err:.2e}", flush=True)
    assert err < 1e-5


def test_gemm_epilogue_bias_relu_mask_accumulate_bf16_out():
    import torch
    from phaser_amd.learned import gemm
    M, N, K = 128, 256, 64
    rng = numpy.random.default_rng(7)
    A = om.round_bf16(rng.normal(size=(M, K)))
    B = om.round_bf16(rng.normal(size=(K, N)))
    bias = rng.normal(size=N).astype(numpy.float32)
    C0 = rng.normal(size=(M, N)).astype(numpy.float32)
    mask = om.round_bf16(rng.normal(size=(M, N)))
    ref = numpy.maximum(A.astype(numpy.float64) @ B + C0 + bias, 0.0) * (mask > 0)
    C = torch.from_numpy(C0).cuda()
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device='cuda')
    gemm(_bf(torch, A), _bf(torch, B), M, N, K, C=C, Cb=Cb, bias=torch.from_numpy(bias).cuda(), relu=True,
         mask=_bf(torch, mask), accumulate=True)
    c = C.cpu().numpy()
    assert rel_err(c, ref) < 1e-5
    numpy.testing.assert_array_equal(Cb.float().cpu().numpy(), om.round_bf16(c))


@pytest.mark.parametrize('M,N,K,epi', [(128, 128, 64, False), (256, 384, 192, True), (2048, 2048, 256, True)])
def test_gemm_nt_double_buffered_with_transposed_copy(M, N, K, epi):
    """phx_gemm_bf16_nt: one K stage and several (the two LDS buffers alternating), the
    128 x 64 tiles (fewer than 256 workgroups of 128 x 128) and the 128 x 128 ones
    (2048 x 2048), with and without the epilogue; Cb and the transposed Ct are the BF16
    rounding of C bit for bit."""
    import torch
    from phaser_amd.learned import gemm_nt
    rng = numpy.random.default_rng(M + N + K)
    A = om.round_bf16(rng.normal(size=(M, K)))
    B = om.round_bf16(rng.normal(size=(N, K)))
    ref = A.astype(numpy.float64) @ B.T.astype(numpy.float64)
    C0 = rng.normal(size=(M, N)).astype(numpy.float32)
    kw = {}
    if epi:
        bias = rng.normal(size=N).astype(numpy.float32)
        mask = om.round_bf16(rng.normal(size=(M, N)))
        ref = numpy.maximum(ref + C0 + bias, 0.0) * (mask > 0)
        kw = dict(bias=torch.from_numpy(bias).cuda(), relu=True, mask=_bf(torch, mask), accumulate=True)
    C = torch.from_numpy(C0).cuda()
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device='cuda')
    Ct = torch.empty((N, M), dtype=torch.bfloat16, device='cuda')
    gemm_nt(_bf(torch, A), _bf(torch, B), M, N, K, C=C, Cb=Cb, Ct=Ct, **kw)
    c = C.cpu().numpy()
    err = rel_err(c, ref)
    print(f"gemm_nt {M}x{N}x{K} epilogue={epi}: rel err {err:.2e}", flush=True)
    assert err < 1e-5
    numpy.testing.assert_array_equal(Cb.float().cpu().numpy(), om.round_bf16(c))
    numpy.testing.assert_array_equal(Ct.float().cpu().numpy(), om.round_bf16(c).T)


def test_transpose_bf16():
    import torch
    from phaser_amd.learned import transpose_bf16
    rng = numpy.random.default_rng(3)
    for rows, cols in ((128, 384), (100, 70), (4096, 2048)):
        A = om.round_bf16(rng.normal(size=(rows, cols)))
        out = torch.empty((cols, rows), dtype=torch.bfloat16, device='cuda')
        transpose_bf16(_bf(torch, A), out)
        numpy.testing.assert_array_equal(out.float().cpu().numpy(), A.T)


def test_gemm_rejects_unsupported_shapes():
    import torch
    from phaser_amd.learned import gemm
    A = torch.zeros((100, 64), dtype=torch.bfloat16, device='cuda')
    B = torch.zeros((64, 128), dtype=torch.bfloat16, device='cuda')
    C = torch.zeros((100, 128), dtype=torch.float32, device='cuda')
    with pytest.raises(NotImplementedError):   # M not a multiple of 128: refused, nothing launched
        gemm(A, B, 100, 128, 64, C=C)


def test_mlp_training_step_matches_oracle():
    import torch
    from phaser_amd.learned import MLP
    sizes = (256, 256, 128, 128)
    batch = 128
    mlp = MLP(sizes, seed=1)
    rng = numpy.random.default_rng(2)
    X = om.round_bf16(rng.normal(size=(batch, sizes[0])))
    T = rng.normal(size=(batch, sizes[-1])).astype(numpy.float32)
    Ws = [w.cpu().numpy().astype(numpy.float64) for w in mlp.W]
    bs = [b.cpu().numpy().astype(numpy.float64) for b in mlp.b]
    Z, acts = om.forward(X, Ws, bs, bf16=True)
    dWs, dbs = om.backward(Z, T.astype(numpy.float64), acts, Ws, bf16=True)
    out = mlp.forward(_bf(torch, X))
    mlp.loss_acc.zero_()
    mlp.backward(torch.from_numpy(T).cuda())
    torch.cuda.synchronize()
    assert rel_err(out.cpu().numpy(), Z) < 2e-3
    assert float(mlp.loss_acc.item()) / Z.size == pytest.approx(om.loss(Z, T), rel=1e-3)
    for l in range(3):
        ew, eb = rel_err(mlp.dW[l].cpu().numpy(), dWs[l]), rel_err(mlp.db[l].cpu().numpy(), dbs[l])
        print(f"layer {l}: dW {ew:.2e} db {eb:.2e}", flush=True)
        assert ew < 1e-2 and eb < 1e-2
    # Adam on the GPU gradients == oracle Adam on the same gradients
    g = [d.cpu().numpy().astype(numpy.float64) for d in mlp.dW]
    w_ref = [w.copy() for w in Ws]
    for l in range(3):
        om.adam(w_ref[l], g[l], numpy.zeros_like(g[l]), numpy.zeros_like(g[l]), 1, lr=1e-3)
    mlp.step(1e-3)
    for l in range(3):
        assert rel_err(mlp.W[l].cpu().numpy(), w_ref[l]) < 1e-6
        numpy.testing.assert_array_equal(mlp.Wb[l].float().cpu().numpy(), om.round_bf16(mlp.W[l].cpu().numpy()))
        numpy.testing.assert_array_equal(mlp.Wbt[l].float().cpu().numpy(), om.round_bf16(mlp.W[l].cpu().numpy()).T)


def test_features_patches_and_stitch():
    import torch
    from phaser_amd.learned import features, patch_targets, stitch
    rng = numpy.random.default_rng(3)
    n = 16
    pats = rng.uniform(0, 50, size=(5, n, n)).astype(numpy.float32)
    f = features(torch.from_numpy(pats).cuda(), n, inv_scale=0.5).float().cpu().numpy()
    ref = om.round_bf16(om.features(pats, 0.5))
    assert numpy.max(numpy.abs(f - ref) / numpy.abs(ref)) <= 2**-7   # at most one bf16 ulp
    # 3 x 3 binning (192^2 patterns -> 64^2 inputs) on a 24 x 24 pattern
    p24 = rng.uniform(0, 50, size=(3, 24, 24)).astype(numpy.float32)
    f3 = features(torch.from_numpy(p24).cuda(), 24, inv_scale=0.5, bin=3).float().cpu().numpy()
    ref3 = om.round_bf16(om.features(p24, 0.5, bin=3))
    assert f3.shape == (3, 64) and numpy.max(numpy.abs(f3 - ref3) / numpy.abs(ref3)) <= 2**-7
    # windows tiling a 2-slice object: extracting then stitching returns the projected
    # object's phase split equally between the slices
    S, H, W, ps = 2, 40, 48, 8
    obj = (numpy.exp(1j * rng.uniform(-0.4, 0.4, (S, H, W))) * rng.uniform(0.9, 1.1, (S, H, W))).astype(numpy.complex64)
    starts = numpy.array([(y, x) for y in range(0, H - ps + 1, 4) for x in range(0, W - ps + 1, 4)], dtype=numpy.int32)
    sd = torch.from_numpy(starts).cuda()
    T = patch_targets(torch.from_numpy(obj).cuda(), sd, ps)
    proj = numpy.prod(obj.astype(numpy.complex128), axis=0)
    t = T.cpu().numpy()
    y0, x0 = starts[3]
    numpy.testing.assert_allclose(t[3, :ps * ps].reshape(ps, ps), numpy.angle(proj[y0:y0 + ps, x0:x0 + ps]), atol=2e-6)
    numpy.testing.assert_allclose(t[3, ps * ps:].reshape(ps, ps), numpy.abs(proj[y0:y0 + ps, x0:x0 + ps]), rtol=2e-6)
    back = stitch(T, sd, S, H, W, ps).cpu().numpy()
    numpy.testing.assert_allclose(numpy.prod(back.astype(numpy.complex128), axis=0), proj, rtol=1e-5, atol=1e-5)
    numpy.testing.assert_allclose(back[0], back[1])


def test_training_lowers_loss_on_a_learnable_map():
    import torch
    from phaser_amd.learned import MLP, train
    rng = numpy.random.default_rng(4)
    rows, din, dout = 2048, 256, 128
    Wt = rng.normal(size=(din, dout)) / numpy.sqrt(din)
    Xn = om.round_bf16(rng.normal(size=(rows, din)))
    Tn = numpy.tanh(Xn @ Wt).astype(numpy.float32)
    mlp = MLP((din, 256, 256, dout), seed=5)
    hist = train(mlp, _bf(torch, Xn), torch.from_numpy(Tn).cuda(), epochs=30, batch=256, lr=1e-3)
    print("loss", [round(h, 5) for h in hist[::5]], flush=True)
    assert hist[-1] < 0.2 * hist[0]
