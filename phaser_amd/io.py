"""Reconstruction-state checkpoints in the layout of phaser/utils/io.py:93-245.

The state is written as a tree of named arrays, exactly the reference's layout:

  type = 'phaser_state', version = '0.1', wavelength
  probe/{data (modes, y, x) complex, sampling (2), extent (2)}
  object/{data (z, y, x) complex, thicknesses (z), zs (z), sampling, extent, corner,
          region_min, region_max (absent when unset)}
  scan (..., 2) f64, tilt (..., 2) f64 (when set)
  iter/{engine_num, engine_iter, total_iter} u64
  progress/{iters u64, detector_errors f64}

Two containers: HDF5 (`.h5`/`.hdf5`, through h5py when it is installed: the files are
then the reference's own) and a numpy `.npz` archive with the same paths as keys
('probe/data', ...), which needs nothing beyond numpy.  h5py is not a dependency of this
package; asking for HDF5 without it raises ImportError naming the alternative.
"""
from pathlib import Path

import numpy

from .sampling import ObjectSampling, Sampling
from .state import IterState, ObjectState, ProbeState, ProgressState, ReconsState

STATE_VERSION = '0.1'


def _np(a):
    if a is None:
        return None
    if hasattr(a, 'detach'):
        return a.detach().cpu().numpy()
    return numpy.asarray(a)


def state_tree(state):
    """The reference's checkpoint layout as {path: value} (io.py:176-245)."""
    t = {'type': 'phaser_state', 'version': STATE_VERSION, 'wavelength': numpy.float64(state.wavelength)}
    if state.probe is not None:
        t['probe/data'] = _np(state.probe.data)
        t['probe/sampling'] = numpy.asarray(state.probe.sampling.sampling, dtype=numpy.float64)
        t['probe/extent'] = numpy.asarray(state.probe.sampling.extent, dtype=numpy.float64)
    if state.object is not None:
        o = state.object
        th = numpy.asarray(_np(o.thicknesses))
        t['object/thicknesses'] = th
        t['object/zs'] = numpy.asarray(o.zs())
        t['object/data'] = _np(o.data)
        t['object/sampling'] = numpy.asarray(o.sampling.sampling, dtype=numpy.float64)
        t['object/extent'] = numpy.asarray(o.sampling.extent, dtype=numpy.float64)
        t['object/corner'] = numpy.asarray(o.sampling.corner, dtype=numpy.float64)
        for k in ('region_min', 'region_max'):
            v = getattr(o.sampling, k)
            if v is not None:
                t[f'object/{k}'] = numpy.asarray(v, dtype=numpy.float64)
    if state.scan is not None:
        t['scan'] = numpy.asarray(_np(state.scan), dtype=numpy.float64)
    if state.tilt is not None:
        t['tilt'] = numpy.asarray(_np(state.tilt), dtype=numpy.float64)
    if state.iter is not None:
        for k in ('engine_num', 'engine_iter', 'total_iter'):
            t[f'iter/{k}'] = numpy.uint64(getattr(state.iter, k))
    if state.progress is not None:
        t['progress/iters'] = numpy.asarray(state.progress.iters).astype(numpy.uint64)
        t['progress/detector_errors'] = numpy.asarray(state.progress.detector_errors).astype(numpy.float64)
    return t


def state_from_tree(t):
    """ReconsState from the layout (io.py:93-174); absent parts are None (a partial state)."""
    def get(k):
        return t[k] if k in t else None
    ty = str(get('type'))
    if ty != 'phaser_state':
        raise ValueError(f"Expected a file of type 'phaser_state', instead got type '{ty}'")
    version = tuple(int(v) for v in str(get('version')).split('.'))
    if version > (0, 1):
        raise ValueError(f"Unsupported file version '{get('version')}'. Maximum supported version is '0.1'.")
    probe = obj = None
    if 'probe/data' in t:
        data = numpy.asarray(t['probe/data'])
        ext = numpy.asarray(t['probe/extent'])
        probe = ProbeState(Sampling(data.shape[-2:], extent=(ext[0], ext[1])), data)
    if 'object/data' in t:
        data = numpy.asarray(t['object/data'])
        samp = ObjectSampling(data.shape[-2:], numpy.asarray(t['object/sampling']), numpy.asarray(t['object/corner']),
                              get('object/region_min'), get('object/region_max'))
        obj = ObjectState(samp, data, numpy.asarray(t['object/thicknesses']))
    it = IterState(*(int(t[f'iter/{k}']) for k in ('engine_num', 'engine_iter', 'total_iter'))) \
        if 'iter/engine_num' in t else IterState()
    prog = ProgressState(iters=numpy.asarray(t['progress/iters']),
                         detector_errors=numpy.asarray(t['progress/detector_errors'])) \
        if 'progress/iters' in t else ProgressState()
    wl = float(t['wavelength']) if 'wavelength' in t else None
    return ReconsState(iter=it, wavelength=wl, probe=probe, object=obj, scan=get('scan'), tilt=get('tilt'),
                       progress=prog)


def _is_hdf5(path):
    return Path(path).suffix.lower() in ('.h5', '.hdf5')


def _h5py():
    try:
        import h5py
    except ImportError:
        raise ImportError("HDF5 checkpoints need h5py, which is not installed; write a '.npz' checkpoint "
                          "(same layout) instead") from None
    return h5py


def write_state(state, path):
    """Write a checkpoint (HDF5 for .h5/.hdf5, else a .npz archive)."""
    path = Path(path)
    tree = state_tree(state)
    if _is_hdf5(path):
        h5py = _h5py()
        with h5py.File(path, 'w') as f:
            for k, v in tree.items():
                if isinstance(v, str):
                    f.create_dataset(k, (), h5py.string_dtype(), v)
                else:
                    f.create_dataset(k, data=v)
        return path
    numpy.savez(path, **{k: numpy.asarray(v) for k, v in tree.items()})
    return path if path.suffix == '.npz' else path.with_suffix(path.suffix + '.npz')


def read_state(path):
    """Read a checkpoint written by write_state (or, with h5py, by phaser itself)."""
    path = Path(path)
    if _is_hdf5(path):
        h5py = _h5py()
        tree = {}
        with h5py.File(path, 'r') as f:
            def visit(name, obj):
                if isinstance(obj, h5py.Dataset):
                    v = obj[()]
                    tree[name] = v.decode() if isinstance(v, bytes) else v
            f.visititems(visit)
        return state_from_tree(tree)
    with numpy.load(path, allow_pickle=False) as f:
        tree = {k: (str(f[k]) if f[k].dtype.kind in 'US' else f[k]) for k in f.files}
    return state_from_tree(tree)
