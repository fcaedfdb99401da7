"""Seeded synthetic inputs for the benchmark configurations: Gaussian-atom phase objects,
the 300 kV / 20 mrad / 100 A-defocus aperture probe with orthogonalised perturbed modes,
jittered raster scans, and Poisson-noised diffraction patterns.

Objects are generated directly on the reconstruction grid that ObjectSampling.from_scan
gives for the scan (pad = n/2 dx + 5 dx, phaser/execute.py:271-273), so the ground truth
and the reconstruction share pixels and the NRMSE needs no resampling.  The atom count
keeps the stated density (e.g. 400 atoms per 256^2 pixels).
"""
from dataclasses import dataclass
import typing as t

import numpy

from .groups import create_rng
from .optics import electron_wavelength, make_focused_probe
from .sampling import ObjectSampling, Sampling
from .state import IterState, ObjectState, Patterns, ProbeState, ProgressState, ReconsState


@dataclass
class SyntheticConfig:
    n: int = 64                  # probe / pattern side (pixels)
    n_modes: int = 1
    n_slices: int = 1
    slice_thickness: float = 20.0  # angstrom
    scan_shape: t.Tuple[int, int] = (16, 16)
    scan_step_px: float = 4.0
    jitter_px: float = 0.5
    dx: float = 0.2              # angstrom / pixel
    energy_ev: float = 300e3
    aperture_mrad: float = 20.0
    defocus: float = 100.0       # angstrom
    atoms_per_256sq: float = 400.0
    atom_sigma_px: float = 1.5
    atom_peak_rad: float = 0.3
    dose: float = 1e5            # electrons per pattern
    base_mode_power: float = 0.7
    seed: int = 0


CONFIG_1 = SyntheticConfig()
CONFIG_2 = SyntheticConfig(n=128, n_modes=4, n_slices=8, scan_shape=(64, 64), scan_step_px=6.0, jitter_px=0.0)
# 24 slices, 192x192 probe x 8 modes, 192 x 192 positions at 7 px, Poisson at 2e4
LSQML_192X8_S24 = SyntheticConfig(n=192, n_modes=8, n_slices=24, scan_shape=(192, 192), scan_step_px=7.0,
                                  jitter_px=0.0, dose=2e4)
# 8 slices, 256x256 probe x 4 modes (the split-pass plane size), 32 x 32 positions at 12 px
LSQML_256X4_S8 = SyntheticConfig(n=256, n_modes=4, n_slices=8, scan_shape=(32, 32), scan_step_px=12.0,
                                 jitter_px=0.0)


def raster_scan(cfg: SyntheticConfig):
    """Raster scan (phaser/utils/scan.py:23-50) in float32 plus seeded uniform jitter."""
    sy, sx = cfg.scan_shape
    dtype = numpy.float32
    yy = numpy.arange(sy, dtype=dtype) - numpy.array(sy / 2., dtype=dtype)
    xx = numpy.arange(sx, dtype=dtype) - numpy.array(sx / 2., dtype=dtype)
    pts = numpy.stack(numpy.meshgrid(yy, xx, indexing='ij'), axis=-1)
    step = cfg.scan_step_px * cfg.dx
    scan = pts * numpy.broadcast_to(numpy.array(step), (2,)).astype(dtype)
    if cfg.jitter_px:
        rng = create_rng(cfg.seed, 'scan_jitter')
        scan = scan + rng.uniform(-cfg.jitter_px, cfg.jitter_px, size=scan.shape).astype(dtype) * numpy.float32(cfg.dx)
    return scan.astype(dtype)


def gaussian_atom_phase(shape, n_atoms, sigma_px, peak, rng):
    """Sum of isotropic Gaussians at uniform positions (phase map, radians)."""
    h, w = shape
    ys = rng.uniform(0, h, n_atoms)
    xs = rng.uniform(0, w, n_atoms)
    yy = numpy.arange(h)[:, None]
    xx = numpy.arange(w)[None, :]
    phase = numpy.zeros(shape, dtype=numpy.float64)
    r = int(numpy.ceil(5 * sigma_px))
    for y0, x0 in zip(ys, xs):
        iy0, iy1 = max(0, int(y0) - r), min(h, int(y0) + r + 2)
        ix0, ix1 = max(0, int(x0) - r), min(w, int(x0) + r + 2)
        phase[iy0:iy1, ix0:ix1] += peak * numpy.exp(
            -((yy[iy0:iy1] - y0)**2 + (xx[:, ix0:ix1] - x0)**2) / (2 * sigma_px**2))
    return phase


def probe_modes(cfg: SyntheticConfig, sampling: Sampling, wavelength: float):
    """Base aperture probe plus orthogonalised seeded perturbations, base power 0.7."""
    ky, kx = sampling.recip_grid(dtype=numpy.float32)
    base = make_focused_probe(ky, kx, wavelength, cfg.aperture_mrad, defocus=cfg.defocus).astype(numpy.complex128)
    if cfg.n_modes == 1:
        return base[None].astype(numpy.complex64)
    rng = create_rng(cfg.seed, 'probe_modes')
    n = cfg.n
    modes = [base / numpy.linalg.norm(base)]
    yy, xx = numpy.meshgrid(numpy.linspace(-1, 1, n), numpy.linspace(-1, 1, n), indexing='ij')
    while len(modes) < cfg.n_modes:
        c = rng.normal(size=6)
        field = (c[0] * yy + c[1] * xx + c[2] * yy * xx + c[3] * (yy**2 - xx**2)
                 + 1j * (c[4] * yy + c[5] * xx))
        cand = base * field
        for m in modes:
            cand = cand - m * numpy.vdot(m, cand)
        cand /= numpy.linalg.norm(cand)
        modes.append(cand)
    powers = numpy.concatenate([[cfg.base_mode_power],
                                numpy.full(cfg.n_modes - 1, (1 - cfg.base_mode_power) / (cfg.n_modes - 1))])
    return (numpy.stack(modes) * numpy.sqrt(powers)[:, None, None]).astype(numpy.complex64)


def make_ground_truth(cfg: SyntheticConfig):
    """(truth object (S, H, W) complex64, object sampling, scan, probe modes, wavelength)."""
    wavelength = electron_wavelength(cfg.energy_ev)
    psamp = Sampling((cfg.n, cfg.n), sampling=(cfg.dx, cfg.dx))
    scan = raster_scan(cfg)
    osamp = ObjectSampling.from_scan(scan, psamp.sampling, psamp.extent / 2. + 5.0 * psamp.sampling)
    H, W = (int(v) for v in osamp.shape)
    rng = create_rng(cfg.seed, 'ground_truth')
    n_atoms = int(round(cfg.atoms_per_256sq * H * W / 256.**2))
    slices = []
    for _s in range(cfg.n_slices):
        ph = gaussian_atom_phase((H, W), max(1, n_atoms // cfg.n_slices if cfg.n_slices > 1 else n_atoms),
                                 cfg.atom_sigma_px, cfg.atom_peak_rad, rng)
        slices.append(numpy.exp(1j * ph))
    truth = numpy.stack(slices).astype(numpy.complex64)
    probe = probe_modes(cfg, psamp, wavelength)
    return truth, osamp, scan, probe, wavelength, psamp


def thicknesses(cfg: SyntheticConfig):
    if cfg.n_slices == 1:
        return numpy.array([], dtype=numpy.float32)
    return numpy.full(cfg.n_slices, cfg.slice_thickness, dtype=numpy.float32)


def add_poisson_noise(intensity: numpy.ndarray, cfg: SyntheticConfig) -> numpy.ndarray:
    """Scale unit-intensity patterns to the dose and draw Poisson counts (seeded), like the
    reference's `poisson` post-load hook (phaser/hooks/preprocessing.py:42-63) without its
    default extra Gaussian."""
    rng = create_rng(cfg.seed, 'poisson_noise')
    return rng.poisson(numpy.maximum(intensity, 0) * cfg.dose).astype(numpy.float32)


def initial_state(cfg: SyntheticConfig, osamp, scan, probe, wavelength, psamp, seed_obj=None):
    """Flat-start reconstruction state: object = 1 + tiny random phase (random_phase_object,
    sigma 1e-6, phaser/utils/object.py:34-53), probe = the known probe modes."""
    rng = create_rng(cfg.seed if seed_obj is None else seed_obj, 'random_phase_object')
    shape = (cfg.n_slices, int(osamp.shape[0]), int(osamp.shape[1]))
    ang = rng.normal(0., 1e-6, shape).astype(numpy.float32)
    obj = (numpy.cos(ang) + numpy.sin(ang) * 1.j).astype(numpy.complex64)
    return ReconsState(
        iter=IterState(), wavelength=wavelength, probe=ProbeState(psamp, probe.copy()),
        object=ObjectState(osamp, obj, thicknesses(cfg)), scan=scan.copy(), tilt=None,
        progress=ProgressState())


def simulate_patterns_gpu(truth, osamp, scan, probe, wavelength, psamp, thick, bwlim_frac=2 / 3, batch=64):
    """Noise-free model intensities (P, n, n) float32 of the truth object, computed with the
    HIP forward kernels (shift, multislice, exit intensity)."""
    import torch
    from .device import GroupWorkspace, stream_handle
    from . import _lib
    from .optics import make_propagators
    state = ReconsState(iter=IterState(), wavelength=wavelength, probe=ProbeState(psamp, probe),
                        object=ObjectState(osamp, truth, thick), scan=scan).to_device()
    n = probe.shape[-1]
    P = int(numpy.prod(scan.shape[:-1]))
    G = (P + batch - 1) // batch
    props = make_propagators(psamp, wavelength, thick, bwlim_frac)
    dummy = numpy.zeros((*scan.shape[:-1], n, n), dtype=numpy.float32)
    ws = GroupWorkspace(state, dummy, numpy.ones((n, n), numpy.float32), props, noise_offset=0.0, noise_eps=1e-3,
                        n_groups=G, bmax=batch)
    ws.set_positions(osamp, scan)
    idx = numpy.arange(P)
    groups = [numpy.array(numpy.unravel_index(idx[g * batch:(g + 1) * batch], scan.shape[:-1])) for g in range(G)]
    ws.set_groups(groups)
    out = torch.empty((P, n * n), dtype=torch.float32, device='cuda')
    for g in range(G):
        d, st = ws.desc_refs[g], stream_handle()
        _lib.call("phx_group_begin", d, st)
        ws.fill_tilt(g, st)
        _lib.call("phx_probe_spectra", d, st)
        _lib.call("phx_shift_probes", d, 0, int(ws.S > 1), st)
        for i in range(1, ws.S - 1):
            _lib.call("phx_forward_slice", d, i, 0, st)
        _lib.call("phx_exit_intensity", d, ctypes_ptr(out), st)
    res = out.cpu().numpy().reshape(*scan.shape[:-1], n, n)
    del ws
    return res


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def make_dataset(cfg: SyntheticConfig, simulate=None):
    """Full synthetic problem: (initial state, Patterns, truth).  `simulate` maps
    (truth, osamp, scan, probe, wavelength, psamp, thicknesses) -> noise-free intensities;
    defaults to the GPU forward model."""
    truth, osamp, scan, probe, wavelength, psamp = make_ground_truth(cfg)
    thick = thicknesses(cfg)
    sim = simulate or simulate_patterns_gpu
    inten = sim(truth, osamp, scan, probe, wavelength, psamp, thick)
    pats = add_poisson_noise(inten, cfg)
    mask = numpy.ones((cfg.n, cfg.n), dtype=numpy.float32)
    state = initial_state(cfg, osamp, scan, probe, wavelength, psamp)
    return state, Patterns(pats, mask), truth
