"""Plan hooks: raw-data loaders, pre/post-processing, and the initial probe, scan, tilt
and object (reference phaser/hooks/__init__.py:19-185 and the modules it names).

These run once on the host before an engine starts (they are not the per-group hot
path), on numpy arrays, with the reference's names, arguments and defaults, so the
`raw_data`, `post_load`, `init` and `post_init` entries of a plan YAML mean the same
here.  Each hook is `f(args, props)`; `resolve(spec, registry)` turns a plan entry
(`{'type': name, **props}` or a bare name) into a callable of `args`.

Known hooks:
  raw_data   empad (EMPAD 4D .raw + metadata .json), npz (patterns + metadata in one
             file), synthetic (the built-in Gaussian-atom generator)
  post_load  crop_data, poisson, scale
  probe      focused            scan  raster          object  random
  tilt       global, custom     post_init  drop_nans, diffraction_align
"""
import json
import logging
import math
import re
from pathlib import Path

import numpy

from .groups import create_rng, create_sparse_groupings
from .optics import electron_wavelength, make_focused_probe
from .sampling import Sampling
from .state import ObjectState, ProbeState

logger = logging.getLogger(__name__)


def resolve(spec, registry, what):
    """A plan hook entry -> callable(args).  `spec` is a name, a mapping with 'type', or
    a callable already."""
    if spec is None:
        return None
    if callable(spec):
        return spec
    d = {'type': spec} if isinstance(spec, str) else dict(spec)
    kind = d.pop('type', None)
    if kind not in registry:
        raise ValueError(f"Unknown {what} hook {kind!r} (known: {', '.join(sorted(registry))})")
    fn = registry[kind]
    return lambda args: fn(args, d)


# ------------------------------------------------------------------------ raw data
def load_4d_empad(path, scan_shape=None, memmap=False):
    """EMPAD raw frames (phaser/io/empad.py:115-159): float32 frames of 130 x 128 (two junk
    rows), frame order scan x fastest; returns (n_y, n_x, 128, 128) with reciprocal y
    flipped.  Without `scan_shape` the shape is read from a 'x{nx}_y{ny}' file name."""
    path = Path(path)
    if scan_shape is None:
        m = re.search(r"x(\d+)_y(\d+)", path.name)
        if not m:
            raise ValueError(f"Unable to infer probe dimensions from name {path.name}")
        n_x, n_y = int(m[1]), int(m[2])
    else:
        n_y, n_x = scan_shape
    a = numpy.memmap(path, dtype=numpy.float32, mode='r') if memmap else numpy.fromfile(path, dtype=numpy.float32)
    if a.size % (130 * 128):
        raise ValueError(f"File not divisible by 130x128 (size={a.size}).")
    a = a.reshape(-1, 130, 128)
    if a.shape[0] != n_x * n_y:
        raise ValueError(f"Got {a.shape[0]} probes, expected {n_x}x{n_y} = {n_x * n_y}.")
    a = a.reshape(n_y, n_x, 130, 128)
    return a[..., 127::-1, :]


def save_4d_empad(arr, path):
    """Write (n_y, n_x, 128, 128) patterns in the EMPAD raw layout (phaser/io/empad.py:
    162-212): two zero rows per frame, reciprocal y flipped."""
    arr = numpy.asarray(arr)
    if arr.ndim != 4 or arr.shape[2:] != (128, 128):
        raise ValueError("Invalid data format")
    out = numpy.zeros((*arr.shape[:2], 130, 128), dtype=numpy.float32)
    out[..., 127::-1, :] = arr.astype(numpy.float32)
    out.tofile(str(path))


def load_empad(args, props):
    """RawDataHook 'empad' (phaser/hooks/io/empad.py:15-86)."""
    path = Path(props['path']).expanduser()
    kv, diff_step, adu = props.get('kv'), props.get('diff_step'), props.get('adu')
    if path.suffix.lower() == '.json':
        meta = json.loads(path.read_text())
        raw = path.parent / meta['raw_filename']
        voltage = kv * 1e3 if kv is not None else meta['voltage']
        diff_step = diff_step or meta.get('diff_step')
        scan_shape = tuple(reversed(meta['scan_shape']))             # [x, y] -> [y, x]
        adu = adu or meta.get('adu')
        needs_scale = meta.get('file_type', 'empad_metadata') != 'pyMultislicer_metadata'
        defocus = meta.get('defocus')
        probe_hook = {'type': 'focused', 'conv_angle': meta.get('conv_angle'),
                      'defocus': defocus * 1e10 if defocus is not None else None}
        corr = meta.get('scan_correction')
        scan_hook = {'type': 'raster', 'shape': scan_shape,
                     'step_size': tuple(s * 1e10 for s in reversed(meta['scan_step'])),   # m -> A
                     'affine': numpy.asarray(corr)[::-1, ::-1] if corr is not None else None,
                     'rotation': -meta['scan_rotation']}
        path = raw
    else:
        voltage = kv * 1e3 if kv is not None else None
        scan_shape = None
        probe_hook = scan_hook = None
        needs_scale = False
    if voltage is None:
        raise ValueError("'kv'/'voltage' must be specified by metadata or passed to 'raw_data'")
    if diff_step is None:
        raise ValueError("'diff_step' must be specified by metadata or passed to 'raw_data'")
    wavelength = electron_wavelength(voltage)
    if not path.exists():
        raise ValueError(f"Couldn't find raw data at path {path}")
    patterns = numpy.fft.ifftshift(load_4d_empad(path, scan_shape, memmap=True), axes=(-1, -2))
    if needs_scale:
        if adu is None:
            logger.warning("ADU not supplied for experimental dataset. This is not recommended.")
        else:
            patterns = patterns / adu
    a = wavelength / (diff_step * 1e-3)
    sampling = Sampling(tuple(patterns.shape[-2:]), extent=(a, a))
    mask = numpy.zeros(patterns.shape[-2:], dtype=numpy.float32)
    mask[2:-2, 2:-2] = 1.
    return {'patterns': patterns, 'mask': numpy.fft.ifftshift(mask, axes=(-1, -2)), 'sampling': sampling,
            'wavelength': wavelength, 'probe_hook': probe_hook, 'scan_hook': scan_hook, 'tilt_hook': None,
            'seed': None}


def load_npz(args, props):
    """Raw data from one .npz: `patterns` (..., n, n) corner-origin, optional `mask`, and
    either `sampling` (2,) or `extent` (2,) of the real-space probe grid, `wavelength` (A)
    or `kv`; optional probe ('conv_angle' mrad, 'defocus' A) and raster scan
    ('scan_step' A) metadata become the probe and scan hooks."""
    path = Path(props['path']).expanduser()
    with numpy.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    pats = d['patterns']
    n = pats.shape[-2:]
    samp = Sampling(n, sampling=tuple(d['sampling'])) if 'sampling' in d else Sampling(n, extent=tuple(d['extent']))
    wl = float(d['wavelength']) if 'wavelength' in d else (
        electron_wavelength(float(d['kv']) * 1e3) if 'kv' in d else None)
    probe_hook = ({'type': 'focused', 'conv_angle': float(d['conv_angle']), 'defocus': float(d.get('defocus', 0.0))}
                  if 'conv_angle' in d else None)
    scan_hook = ({'type': 'raster', 'shape': tuple(int(v) for v in pats.shape[:2]),
                  'step_size': tuple(float(v) for v in numpy.broadcast_to(d['scan_step'], (2,)))}
                 if 'scan_step' in d and pats.ndim == 4 else None)
    return {'patterns': pats, 'mask': d.get('mask', numpy.ones(n, dtype=pats.dtype)), 'sampling': samp,
            'wavelength': wl, 'probe_hook': probe_hook, 'scan_hook': scan_hook, 'tilt_hook': None, 'seed': None}


def load_synthetic(args, props):
    """Raw data from the built-in synthetic generator (phaser_amd.synthetic): props are
    SyntheticConfig fields.  Noise-free intensities times the dose (add a `poisson`
    post_load step for counting noise); the probe and scan hooks are those the data
    were made with."""
    from .synthetic import SyntheticConfig, make_ground_truth, thicknesses, simulate_patterns_gpu
    cfg = SyntheticConfig(**props)
    truth, osamp, scan, probe, wavelength, psamp = make_ground_truth(cfg)
    inten = simulate_patterns_gpu(truth, osamp, scan, probe, wavelength, psamp, thicknesses(cfg))
    pats = (numpy.asarray(inten, dtype=numpy.float64) * cfg.dose).astype(numpy.float32)
    return {'patterns': pats, 'mask': numpy.ones((cfg.n, cfg.n), numpy.float32), 'sampling': psamp,
            'wavelength': wavelength,
            'probe_hook': {'type': 'focused', 'conv_angle': cfg.aperture_mrad, 'defocus': cfg.defocus},
            'scan_hook': {'type': 'raster', 'shape': tuple(cfg.scan_shape),
                          'step_size': cfg.scan_step_px * cfg.dx},
            'tilt_hook': None, 'seed': cfg.seed}


RAW_DATA = {'empad': load_empad, 'npz': load_npz, 'synthetic': load_synthetic}


# ----------------------------------------------------------------------- post_load
def crop_data(raw, props):
    """hooks/preprocessing.py:18-35."""
    if raw['patterns'].ndim != 4:
        raise ValueError(f"'crop_data' expects a 4D array of patterns, got shape {raw['patterns'].shape} instead")
    y_i, y_f, x_i, x_f = props['crop']
    raw['patterns'] = raw['patterns'][slice(y_i, y_f), slice(x_i, x_f)]
    sh = raw.get('scan_hook')
    if sh is not None and sh.get('type') == 'raster':
        raw['scan_hook'] = dict(sh, shape=raw['patterns'].shape[:2])
    return raw


def scale_patterns(raw, props):
    """hooks/preprocessing.py:38-40."""
    raw['patterns'] = raw['patterns'] * props['scale']
    return raw


def add_poisson_noise(raw, props):
    """hooks/preprocessing.py:43-63: optional scale, Poisson counts from the seeded
    'poisson_noise' stream, plus N(0, gaussian) (default 1e-3)."""
    dtype = raw['patterns'].dtype
    scale = props.get('scale')
    gaussian = props.get('gaussian', 1.0e-3)
    pats = numpy.asarray(raw['patterns'])
    if scale is not None:
        pats = pats * scale
    rng = create_rng(raw.get('seed', None), 'poisson_noise')
    out = rng.poisson(pats).astype(dtype)
    if gaussian is not None:
        out += rng.normal(scale=gaussian, size=out.shape)
    raw['patterns'] = out
    return raw


POST_LOAD = {'crop_data': crop_data, 'poisson': add_poisson_noise, 'scale': scale_patterns}


# ------------------------------------------------------------------------ init hooks
def focused_probe(args, props):
    """ProbeHook 'focused' (hooks/probe.py:8-24)."""
    if props.get('conv_angle') is None:
        raise ValueError("Probe 'conv_angle' must be specified by metadata or manually")
    if props.get('defocus') is None:
        raise ValueError("Probe 'defocus' must be specified by metadata or manually")
    samp = args['sampling']
    ky, kx = samp.recip_grid(dtype=args['dtype'])
    probe = make_focused_probe(ky, kx, args['wavelength'], props['conv_angle'], defocus=props['defocus'])
    return ProbeState(samp, probe)


def make_raster_scan(shape, scan_step, rotation=0., dtype=numpy.float64):
    """utils/scan.py:23-50: (n_y, n_x, 2) positions about the scan centre."""
    yy = numpy.arange(shape[0], dtype=dtype) - numpy.array(shape[0] / 2., dtype=dtype)
    xx = numpy.arange(shape[1], dtype=dtype) - numpy.array(shape[1] / 2., dtype=dtype)
    pts = numpy.stack(numpy.meshgrid(yy, xx, indexing='ij'), axis=-1)
    if rotation != 0.:
        th = rotation * numpy.pi / 180.
        mat = numpy.array([[numpy.cos(th), -numpy.sin(th)], [numpy.sin(th), numpy.cos(th)]], dtype=dtype)
        pts = pts @ mat.T
    return pts * numpy.broadcast_to(numpy.array(scan_step), (2,)).astype(dtype)


def raster_scan(args, props):
    """ScanHook 'raster' (hooks/scan.py:9-27)."""
    if props.get('shape') is None:
        raise ValueError("scan 'shape' must be specified by metadata or manually")
    if props.get('step_size') is None:
        raise ValueError("scan 'step_size' must be specified by metadata or manually")
    scan = make_raster_scan(tuple(props['shape']), props['step_size'], props.get('rotation') or 0.0,
                            dtype=args['dtype'])
    if props.get('affine') is not None:
        scan = scan @ numpy.asarray(props['affine'], dtype=scan.dtype).T
    return scan


def random_phase_object(shape, sigma=1e-6, seed=None, dtype=numpy.complex64):
    """utils/object.py:34-53: exp(i N(0, sigma)) from the seeded 'random_phase_object'
    stream."""
    rng = create_rng(seed, 'random_phase_object')
    real = numpy.float32 if numpy.dtype(dtype) == numpy.complex64 else numpy.float64
    ang = numpy.array(rng.normal(0., sigma, tuple(shape)), dtype=real)
    return numpy.cos(ang) + numpy.sin(ang) * 1.j


def random_object(args, props):
    """ObjectHook 'random' (hooks/object.py:8-25)."""
    samp = args['sampling']
    cdt = numpy.complex64 if numpy.dtype(args['dtype']) == numpy.float32 else numpy.complex128
    if args.get('slices') is not None:
        th = numpy.array(args['slices'], dtype=args['dtype'])
        shape = (len(th), *samp.shape)
    else:
        th = numpy.array([], dtype=args['dtype'])
        shape = tuple(samp.shape)
    obj = random_phase_object(shape, props.get('sigma', 1e-6), seed=args.get('seed'), dtype=cdt)
    return ObjectState(samp, obj, th)


def global_tilt(args, props):
    """TiltHook 'global' (hooks/tilt.py:8-23): [ty, tx] mrad at every position."""
    ty, tx = props['tilt']
    return numpy.broadcast_to(numpy.array([ty, tx], dtype=numpy.float32), (*args['shape'], 2)).copy()


def custom_tilt(args, props):
    """TiltHook 'custom' (hooks/tilt.py:26-55): an (ny, nx, 2) or (ny*nx, 2) .npy."""
    path = Path(props['path']).expanduser()
    if not path.exists():
        raise FileNotFoundError(f"Custom tilt file not found: {path}")
    data = numpy.load(path, allow_pickle=False)
    shape = tuple(args['shape'])
    if data.ndim == 3 and data.shape == (*shape, 2):
        return data.astype(numpy.float32)
    if data.ndim == 2 and data.shape == (math.prod(shape), 2):
        return data.reshape(*shape, 2).astype(numpy.float32)
    raise ValueError(f"Loaded tilt data shape {data.shape} does not match the scan shape {shape}")


PROBE = {'focused': focused_probe}
SCAN = {'raster': raster_scan}
OBJECT = {'random': random_object}
TILT = {'global': global_tilt, 'custom': custom_tilt}


# ------------------------------------------------------------------------ post_init
def drop_nan_patterns(args, props):
    """PostInitHook 'drop_nans' (hooks/preprocessing.py:66-104): patterns with more than
    `threshold` NaN pixels are dropped with their scan (and tilt) positions."""
    threshold = props.get('threshold', 0.9)
    data, state = args['data'], args['state']
    scan = numpy.asarray(state.scan).reshape(-1, 2)
    tilt = None if state.tilt is None else numpy.asarray(state.tilt).reshape(-1, 2)
    pats = numpy.asarray(data.patterns).reshape(-1, *data.patterns.shape[-2:])
    frac = numpy.sum(numpy.isnan(pats), axis=(-1, -2)) / math.prod(pats.shape[-2:])
    mask = frac > threshold
    if (n := int(numpy.sum(mask))):
        logger.info(f"Dropping {n}/{pats.shape[0]} patterns which are at least {threshold:.1%} NaN values")
        pats = pats[~mask]
        if scan.shape[0] == mask.size:
            scan = scan[~mask]
        elif scan.shape[0] != pats.shape[0]:
            raise ValueError(f"# of scan positions {scan.shape[0]} doesn't match # of patterns"
                             f" before ({mask.size}) or after ({pats.shape[0]}) filtering")
        if tilt is not None:
            if tilt.shape[0] == mask.size:
                tilt = tilt[~mask]
            elif tilt.shape[0] != pats.shape[0]:
                raise ValueError(f"# of tilt positions {tilt.shape[0]} doesn't match # of patterns"
                                 f" before ({mask.size}) or after ({pats.shape[0]}) filtering")
    state.scan = scan
    state.tilt = tilt
    data.patterns = pats
    return data, state


def _shift_bilinear(arr, shift):
    """out[.., y, x] = arr[.., y + sy, x + sx], linear interpolation with zeros outside
    (an affine transform with unit matrix and offset `shift`, as utils/image.py:151-195)."""
    from .sampling import affine_transform
    arr = numpy.asarray(arr)
    return affine_transform(arr, [1., 1.], shift, tuple(arr.shape[-2:]), order=1, cval=0.0)


def diffraction_align(args, props=None):
    """PostInitHook 'diffraction_align' (hooks/preprocessing.py:107-141): shift every
    pattern (and the mask) so the mean pattern's centre of mass sits at zero frequency."""
    data, state = args['data'], args['state']
    pats = data.patterns
    groups = create_sparse_groupings(pats.shape[:-2], 128)
    dtype = pats.dtype
    sum_pattern = numpy.zeros(pats.shape[-2:], dtype=dtype)
    for g in groups:
        p = numpy.asarray(pats[tuple(g)]) * numpy.asarray(data.pattern_mask)
        sum_pattern += numpy.nansum(p, axis=tuple(range(p.ndim - 2)))
    mean_pattern = sum_pattern / math.prod(pats.shape[:-2])
    ky, kx = Sampling(tuple(mean_pattern.shape), extent=(1.0, 1.0)).recip_grid(dtype=dtype)
    shift = numpy.array([numpy.nansum(ky * mean_pattern), numpy.nansum(kx * mean_pattern)]) / numpy.nansum(mean_pattern)
    logger.info(f"Shifting diffraction patterns by ({shift[1]}, {shift[0]}) px")

    def shifted(a):
        return numpy.fft.ifftshift(_shift_bilinear(numpy.fft.fftshift(a, axes=(-2, -1)), shift), axes=(-2, -1))

    pats = numpy.array(pats)
    for g in groups:
        pats[tuple(g)] = shifted(pats[tuple(g)])
    data.patterns = pats
    data.pattern_mask = shifted(data.pattern_mask)
    return data, state


POST_INIT = {'drop_nans': drop_nan_patterns, 'diffraction_align': diffraction_align}
