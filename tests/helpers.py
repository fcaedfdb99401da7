"""Shared test helpers: oracle/product state conversion and error metrics."""
import numpy

import oracle.num as onum
import oracle.object as oobj
import oracle.state as ost


def rel_err(a, b):
    a = numpy.asarray(a, dtype=numpy.complex128)
    b = numpy.asarray(b, dtype=numpy.complex128)
    return float(numpy.linalg.norm((a - b).ravel()) / max(numpy.linalg.norm(b.ravel()), 1e-300))


def to_oracle_state(state, dtype=numpy.float32):
    """Product ReconsState (numpy arrays) -> oracle ReconsState, optionally in FP64."""
    cdt = numpy.complex64 if dtype == numpy.float32 else numpy.complex128
    ps = state.probe.sampling
    os_ = state.object.sampling
    probe = ost.ProbeState(onum.Sampling(tuple(ps.shape), sampling=tuple(ps.sampling)),
                           numpy.array(numpy.asarray(state.probe.data), dtype=cdt))
    obj = ost.ObjectState(oobj.ObjectSampling(tuple(os_.shape), os_.sampling, os_.corner, os_.region_min, os_.region_max),
                          numpy.array(numpy.asarray(state.object.data), dtype=cdt),
                          numpy.asarray(state.object.thicknesses, dtype=dtype))
    tilt = getattr(state, 'tilt', None)
    return ost.ReconsState(iter=ost.IterState(), wavelength=state.wavelength, probe=probe, object=obj,
                           scan=numpy.array(state.scan, dtype=dtype),
                           tilt=None if tilt is None else numpy.array(tilt, dtype=dtype))


def oracle_simulate(truth, osamp, scan, probe, wavelength, psamp, thick, bwlim_frac=2 / 3):
    """Noise-free exit intensities with the oracle forward model (FP64, small cases)."""
    from oracle.optics import make_propagators, fourier_shift_filter
    samp = onum.Sampling(tuple(psamp.shape), sampling=tuple(psamp.sampling))
    ky, kx = samp.recip_grid(dtype=numpy.float64)
    props = make_propagators(ky, kx, samp.k_max, wavelength, thick, bwlim_frac)
    o = oobj.ObjectSampling(tuple(osamp.shape), osamp.sampling, osamp.corner)
    flat = scan.reshape(-1, 2).astype(numpy.float64)
    n = probe.shape[-1]
    out = numpy.empty((flat.shape[0], n, n))
    P = probe.astype(numpy.complex128)
    for k, pos in enumerate(flat):
        cut = o.get_view_at_pos(truth.astype(numpy.complex128), pos[None], (n, n))[0]
        filt = fourier_shift_filter(ky, kx, o.get_subpx_shifts(pos[None], (n, n)))[0]
        psi = onum.ifft2(onum.fft2(P) * filt)
        for s in range(truth.shape[0] - 1):
            psi = onum.ifft2(onum.fft2(psi * cut[s]) * props[s])
        out[k] = numpy.sum(onum.abs2(onum.fft2(psi * cut[-1])), axis=0)
    return out.reshape(*scan.shape[:-1], n, n).astype(numpy.float32)
