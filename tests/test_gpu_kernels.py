"""GPU parity of the building-block kernels against the oracle: the convention-exact
fft2/ifft2 and the cutout gather / scatter-add / set (reference inputs of
tests/test_object.py)."""
import ctypes

import numpy
import pytest

import oracle.num as onum
import oracle.object as oobj
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _st():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize('n', [16, 32, 64, 128, 192, 256])
def test_fft2_ifft2_match_oracle(n):
    import torch
    from phaser_amd import _lib
    from phaser_amd.optics import twiddles
    rng = numpy.random.default_rng(n)
    a = (rng.normal(size=(3, n, n)) + 1j * rng.normal(size=(3, n, n))).astype(numpy.complex64)
    da = torch.from_numpy(a).cuda()
    out = torch.empty_like(da)
    tw = torch.from_numpy(twiddles(n)).cuda()
    _lib.call('phx_fft2_c64', _ptr(da), _ptr(out), 3, n, 0, _ptr(tw), _st())
    assert rel_err(out.cpu().numpy(), onum.fft2(a.astype(numpy.complex128))) < 2e-6
    _lib.call('phx_fft2_c64', _ptr(da), _ptr(out), 3, n, 1, _ptr(tw), _st())
    assert rel_err(out.cpu().numpy(), onum.ifft2(a.astype(numpy.complex128))) < 2e-6


def test_fft2_closed_forms_on_gpu():
    """tests/test_num.py:102-126 closed forms, on an even 16x16 grid."""
    import torch
    from phaser_amd import _lib
    from phaser_amd.optics import twiddles
    n = 16
    a = numpy.zeros((1, n, n), numpy.complex64)
    a[0, n // 2, n // 2] = 5.0   # centred delta -> constant 5/n
    da = torch.from_numpy(a).cuda()
    out = torch.empty_like(da)
    tw = torch.from_numpy(twiddles(n)).cuda()
    _lib.call('phx_fft2_c64', _ptr(da), _ptr(out), 1, n, 0, _ptr(tw), _st())
    numpy.testing.assert_allclose(out.cpu().numpy()[0], numpy.full((n, n), 5.0 / n), atol=1e-6)
    b = numpy.full((1, n, n), 1 + 1j, numpy.complex64)   # constant -> delta at k=0
    db = torch.from_numpy(b).cuda()
    _lib.call('phx_fft2_c64', _ptr(db), _ptr(out), 1, n, 0, _ptr(tw), _st())
    exp = numpy.zeros((n, n), numpy.complex64)
    exp[0, 0] = (1 + 1j) * n
    numpy.testing.assert_allclose(out.cpu().numpy()[0], exp, atol=1e-5)


RAMP_POS = numpy.array([[0., 0.], [-50., -50.], [-50., 50.], [50., 50.]])


def _starts(samp, pos, shape):
    import torch
    return torch.from_numpy(samp.start_idxs(pos, shape).astype(numpy.int64)).cuda()


@pytest.mark.parametrize('kind', ['float', 'complex'])
def test_get_cutouts_bit_exact(kind):
    """Inputs of tests/test_object.py:147-176; gather is bit-exact against numpy slicing."""
    import torch
    from phaser_amd import _lib
    samp = oobj.ObjectSampling((200, 200), (1.0, 1.0))
    yy, xx = numpy.meshgrid(numpy.arange(200), numpy.arange(200), indexing='ij')
    if kind == 'float':
        obj = (yy + 2. * xx).astype(numpy.float32)
    else:
        obj = numpy.exp((yy + 2. * xx).astype(numpy.complex64) * 1.j / 100).astype(numpy.complex64)
    expected = samp.get_view_at_pos(obj, RAMP_POS, (64, 64))
    dobj = torch.from_numpy(obj).cuda()
    out = torch.empty((4, 64, 64), dtype=dobj.dtype, device='cuda')
    st = _starts(samp, RAMP_POS, (64, 64))
    if kind == 'float':
        _lib.call('phx_get_cutouts_f32', _ptr(dobj), _ptr(out), _ptr(st), 4, 1, 200, 200, 64, 64, 1.0, _st())
    else:
        _lib.call('phx_get_cutouts_c64', _ptr(dobj), _ptr(out), _ptr(st), 4, 1, 200, 200, 64, 64, 1.0, 0.0, _st())
    numpy.testing.assert_array_equal(out.cpu().numpy(), expected)


def test_get_cutouts_multislice_ramp_and_oob_default():
    """tests/test_object.py:266-296 ramp corners (3 slices, 80x100 cutouts) and the cupy
    out-of-bounds default value."""
    import torch
    from phaser_amd import _lib
    samp = oobj.ObjectSampling((200, 200), (1.0, 1.0))
    zz, yy, xx = numpy.indices((3, 200, 200), dtype=numpy.float32)
    obj = 1000. * zz + (2. * yy + xx)
    pos = numpy.array([[-50., -50.], [-50., 50.], [50., -50.], [50., 50.]])
    dobj = torch.from_numpy(obj).cuda()
    out = torch.empty((4, 3, 80, 100), dtype=torch.float32, device='cuda')
    st = _starts(samp, pos, (80, 100))
    _lib.call('phx_get_cutouts_f32', _ptr(dobj), _ptr(out), _ptr(st), 4, 3, 200, 200, 80, 100, 1.0, _st())
    arr = out.cpu().numpy()
    numpy.testing.assert_array_equal(arr[..., 0, 0], [[20., 1020., 2020.], [120., 1120., 2120.],
                                                      [220., 1220., 2220.], [320., 1320., 2320.]])
    numpy.testing.assert_array_equal(numpy.diff(arr, axis=-1), 1.)
    numpy.testing.assert_array_equal(numpy.diff(arr, axis=-2), 2.)
    # a start of (-10, 190) leaves the object on two sides: those pixels read the default
    st2 = torch.tensor([[-10, 190]], dtype=torch.int64, device='cuda')
    out2 = torch.empty((1, 1, 16, 16), dtype=torch.float32, device='cuda')
    _lib.call('phx_get_cutouts_f32', _ptr(dobj), _ptr(out2), _ptr(st2), 1, 1, 200, 200, 16, 16, 1.0, _st())
    o2 = out2.cpu().numpy()[0, 0]
    assert numpy.all(o2[:10] == 1.0) and numpy.all(o2[:, 10:] == 1.0)
    numpy.testing.assert_array_equal(o2[10:, :10], obj[0, 0:6, 190:200])


def test_add_cutouts_float_exact_and_complex():
    """tests/test_object.py:179-244 inputs: 30 cutouts of 10.0 (sum invariant, exact since
    every partial sum is an integer) and of 10*exp(i phase)."""
    import torch
    from phaser_amd import _lib
    from tests.test_oracle_kat import ADD_POS
    samp = oobj.ObjectSampling((200, 200), (1.0, 1.0))
    st = _starts(samp, ADD_POS, (64, 64))
    obj = torch.zeros((200, 200), dtype=torch.float32, device='cuda')
    cut = torch.full((30, 64, 64), 10., dtype=torch.float32, device='cuda')
    _lib.call('phx_add_cutouts_f32', _ptr(obj), _ptr(cut), _ptr(st), 30, 1, 200, 200, 64, 64, _st())
    got = obj.cpu().numpy()
    expected = samp.add_view_at_pos(numpy.zeros((200, 200), numpy.float32), ADD_POS,
                                    numpy.full((30, 64, 64), 10., numpy.float32))
    numpy.testing.assert_array_equal(got, expected)
    assert got.sum() == 10. * 30 * 64 * 64

    phases = numpy.random.default_rng(0).uniform(0, 6.28, 30).astype(numpy.float32)
    c = (numpy.full((30, 64, 64), 10., numpy.complex64) * numpy.exp(1.j * phases[:, None, None])).astype(numpy.complex64)
    cobj = torch.zeros((200, 200), dtype=torch.complex64, device='cuda')
    _lib.call('phx_add_cutouts_c64', _ptr(cobj), _ptr(torch.from_numpy(c).cuda()), _ptr(st), 30, 1, 200, 200, 64, 64, _st())
    exp_c = samp.add_view_at_pos(numpy.zeros((200, 200), numpy.complex64), ADD_POS, c)
    assert rel_err(cobj.cpu().numpy(), exp_c) < 1e-6


def test_set_cutouts_max():
    """tests/test_object.py:299-362: max after set equals the set value."""
    import torch
    from phaser_amd import _lib
    from tests.test_oracle_kat import ADD_POS
    samp = oobj.ObjectSampling((200, 200), (1.0, 1.0))
    st = _starts(samp, ADD_POS, (64, 64))
    obj = torch.zeros((200, 200), dtype=torch.complex64, device='cuda')
    cut = torch.full((30, 64, 64), 10. + 15.j, dtype=torch.complex64, device='cuda')
    _lib.call('phx_set_cutouts_c64', _ptr(obj), _ptr(cut), _ptr(st), 30, 1, 200, 200, 64, 64, _st())
    got = obj.cpu().numpy()
    assert numpy.max(onum.abs2(got)) == onum.abs2(10. + 15.j)
    expected = samp.set_view_at_pos(numpy.zeros((200, 200), numpy.complex64), ADD_POS,
                                    numpy.full((30, 64, 64), 10. + 15.j, numpy.complex64))
    numpy.testing.assert_array_equal(got != 0, expected != 0)


def test_scramble_roundtrip():
    import torch
    from phaser_amd import _lib
    from phaser_amd.optics import scramble
    rng = numpy.random.default_rng(3)
    for n in (16, 32, 64, 128, 192):
        a = rng.normal(size=(2, n, n)).astype(numpy.float32)
        da = torch.from_numpy(a).cuda()
        out = torch.empty_like(da)
        back = torch.empty_like(da)
        _lib.call('phx_scramble', _ptr(da), _ptr(out), 2, n, 0, 0, _st())
        numpy.testing.assert_array_equal(out.cpu().numpy(), scramble(a))
        _lib.call('phx_scramble', _ptr(out), _ptr(back), 2, n, 0, 1, _st())
        numpy.testing.assert_array_equal(back.cpu().numpy(), a)


def test_fft256_roundtrip_and_parseval_at_batch_64():
    """n = 256 (two launches, the column pass in place): ifft2(fft2(a)) = a and
    sum |fft2(a)|^2 = sum |a|^2 (norm='ortho') over 64 planes."""
    import torch
    from phaser_amd import _lib
    from phaser_amd.optics import twiddles
    n, batch = 256, 64
    g = torch.Generator().manual_seed(256)
    a = torch.complex(torch.randn(batch, n, n, generator=g), torch.randn(batch, n, n, generator=g)).cuda()
    f = torch.empty_like(a)
    back = torch.empty_like(a)
    tw = torch.from_numpy(twiddles(n)).cuda()
    _lib.call('phx_fft2_c64', _ptr(a), _ptr(f), batch, n, 0, _ptr(tw), _st())
    _lib.call('phx_fft2_c64', _ptr(f), _ptr(back), batch, n, 1, _ptr(tw), _st())
    torch.cuda.synchronize()
    assert float((back - a).abs().max()) < 1e-5
    ea = a.abs().square().double().sum(dim=(1, 2))
    ef = f.abs().square().double().sum(dim=(1, 2))
    assert float(((ef - ea) / ea).abs().max()) < 1e-6
