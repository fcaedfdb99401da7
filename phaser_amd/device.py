"""Device workspace for the group step: HBM-resident constants, patterns and buffers, one
`phx_group_desc` per group slot, and the launch sequences of the LSQ-ML and ePIE group
steps and their dry runs.

HBM layout (config with P positions, n x n planes, M modes, S slices, H x W slices, B slots):
  object (S, H, W) c64 and probe (M, n, n) c64 are the state tensors themselves;
  patterns_s (P, n*n) f32 and mask_s (n*n) f32 are resident, in scrambled frequency order;
  props_s (S-1, n*n) c64 scrambled; psi (S, B, M, n, n) c64 holds every slice's forward
  wave for the backward pass; chi_a / chi_b (B, M, n, n) ping-pong the back-propagated
  exit-wave update; farfield (B, M, n*n) c64 and intensity (B, n*n, M) f32 carry the exit
  spectra between the two far-field kernels; gpm (S, H, W) f32 (illumination scatter of the
  per-slice entry points, unused by the engines) and d_probe (M, n*n) c64 are accumulators;
  red holds the FP64 step-size sums.
Per-iteration tables: group_pos (G, B) int32, group_size (G,) int32, start (P, 2) int32,
shift (P, 2) f32.  Nothing is allocated after construction, so a whole iteration can be
captured in one HIP graph and replayed.
"""
import ctypes

import numpy

from . import _lib
from .optics import fft_lanes, scr_to_freq, twiddles

INT32_MAX = 2**31 - 1
# largest plane whose LSQ-ML probe step runs in strips (phx_lsqml_probe_apply_strips)
PROBE_STRIPS_MAX_N = 128
# plane sides with group kernels; 256 x 256 planes run every engine step in split passes
# (row and column launches through HBM; the gradient engine without its recompute mode)
PLANE_SIDES = (16, 32, 64, 128, 192, 256)
SPLIT_N = 256
# largest plane side the whole-sweep kernels are used for (larger: one launch per slice)
FUSED_MAX_N = 128
# the object-update launch of group g also clears group g+1's accumulators
# (phx_lsqml_object_update_slices_begin) instead of a separate phx_group_begin
FOLD_GROUP_BEGIN = True


def _require_torch_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the MI355X backend needs a visible GPU (no CPU fallback)")
    return torch


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def host_patterns(patterns):
    """Diffraction patterns as given to an engine: a torch tensor already on the GPU is
    kept there (no host round trip); anything else becomes a numpy array."""
    if hasattr(patterns, 'is_cuda') and patterns.is_cuda:
        return patterns
    return numpy.asarray(patterns)


def to_device_scrambled(torch, arr, n, is_complex, device):
    """Upload (batch, n, n) natural-order frequency data and scramble it on the device
    (a device tensor is scrambled where it is)."""
    if hasattr(arr, 'is_cuda') and arr.is_cuda:
        batch = arr.reshape(-1, n, n).shape[0]
        src = arr.reshape(batch, n * n).contiguous()
    else:
        arr = numpy.ascontiguousarray(arr)
        batch = arr.reshape(-1, n, n).shape[0]
        src = torch.from_numpy(arr.reshape(batch, n * n)).to(device)
    out = torch.empty_like(src)
    st = stream_handle()
    for b0 in range(0, batch, 65535):
        b1 = min(batch, b0 + 65535)
        _lib.call("phx_scramble", ctypes.c_void_p(src[b0:b1].data_ptr()), ctypes.c_void_p(out[b0:b1].data_ptr()),
                  b1 - b0, n, int(is_complex), 0, st)
    del src
    return out


def zero_column_waves(props, n):
    """Bitmask of the plane workgroup's waves whose column-phase columns (scrambled
    positions) are zero in every propagator: the kernels skip their column passes in a
    propagation.  `props` is (S-1, n, n) in natural frequency order, or None."""
    if props is None:
        return 0
    L = fft_lanes(n)
    cols_per_wave = 64 // L
    zero_col = numpy.all(numpy.asarray(props).reshape(-1, n, n) == 0, axis=(0, 1))[scr_to_freq(n)]
    mask = 0
    for w in range(n * L // 64):
        if zero_col[w * cols_per_wave:(w + 1) * cols_per_wave].all():
            mask |= 1 << w
    return mask


def tilt_tables(tilt, thicknesses, n_pos):
    """Host tables of the tilted propagators (tilt_propagators, simulation.py:193-222):
    (tan(1e-3 tilt) as (P, 2) float32 taken in FP64, slice thicknesses dz_j as (S-1)
    float32), or (None, None) when there is nothing to tilt (no tilts, all-zero tilts,
    whose ramps are exactly 1, or a single slice)."""
    if tilt is not None and hasattr(tilt, 'detach'):
        tilt = tilt.detach().cpu().numpy()
    dz = numpy.asarray(thicknesses, dtype=numpy.float64)[:-1]
    if tilt is None or len(dz) == 0 or not numpy.any(numpy.asarray(tilt) != 0):
        return None, None
    tilt = numpy.asarray(tilt, dtype=numpy.float64).reshape(-1, 2)
    if tilt.shape[0] != n_pos:
        raise ValueError(f"tilt has {tilt.shape[0]} positions, the scan {n_pos}")
    tan = numpy.ascontiguousarray(numpy.tan(tilt * 1e-3).astype(numpy.float32))
    return tan, numpy.ascontiguousarray(dz.astype(numpy.float32))


class GroupWorkspace:
    """Buffers and descriptors for one engine run over a fixed set of positions."""

    def __init__(self, state, patterns, pattern_mask, props, *, noise_offset, noise_eps, n_groups, bmax,
                 noise_var=1.1, psi_slots=0, gradients=False, keep_chi=False, fused=True, device='cuda'):
        torch = _require_torch_gpu()
        _lib.load()
        self.torch = torch
        self.device = device
        obj, probe = state.object.data, state.probe.data
        assert obj.dtype == torch.complex64 and probe.dtype == torch.complex64
        assert obj.is_contiguous() and probe.is_contiguous()
        self.S, self.H, self.W = (int(v) for v in obj.shape)
        self.M, ny, nx = (int(v) for v in probe.shape)
        if ny != nx or ny not in PLANE_SIDES:
            raise NotImplementedError(f"probe shape {ny}x{nx}: the plane kernels support square planes of side "
                                      "16, 32, 64, 128, 192 or 256")
        self.n = n = ny
        self.NN = n * n
        pats = host_patterns(patterns)
        if hasattr(pats, 'is_cuda') and pats.dtype != torch.float32:
            raise TypeError("device patterns must be float32")
        self.scan_shape = tuple(int(v) for v in pats.shape[:-2])
        self.P = int(numpy.prod(self.scan_shape))
        self.G, self.B = int(n_groups), int(bmax)
        f32, c64, i32 = torch.float32, torch.complex64, torch.int32
        dev = device

        # constants
        ky, kx = state.probe.sampling.recip_1d(numpy.float32)
        f = scr_to_freq(n)
        self.ky_s = torch.from_numpy(numpy.ascontiguousarray(ky[f])).to(dev)
        self.kx_s = torch.from_numpy(numpy.ascontiguousarray(kx[f])).to(dev)
        self.twiddle = torch.from_numpy(twiddles(n)).to(dev)
        self.patterns_s = to_device_scrambled(torch, pats if hasattr(pats, 'is_cuda') else pats.astype(numpy.float32),
                                              n, False, dev)
        self.mask_s = to_device_scrambled(torch, numpy.asarray(pattern_mask, dtype=numpy.float32), n, False, dev).reshape(-1)
        self.props_s = None if props is None else to_device_scrambled(torch, props.astype(numpy.complex64), n, True, dev)
        self.prop_zero_waves = zero_column_waves(props, n)
        # [beta_obj, beta_probe, reg_obj, reg_probe, gamma, offset, eps, var, loss_scale]
        self.params = torch.zeros(16, dtype=f32, device=dev)
        self.params_host = numpy.zeros(16, dtype=numpy.float32)
        self.params_host[5] = noise_offset
        self.params_host[6] = noise_eps
        self.params_host[7] = noise_var
        self.params_host[8] = 1.0
        self.params.copy_(torch.from_numpy(self.params_host))

        # per-iteration tables
        self.group_pos = torch.zeros((self.G, self.B), dtype=i32, device=dev)
        self.group_size = torch.zeros(self.G, dtype=i32, device=dev)
        self.start = torch.zeros((self.P, 2), dtype=i32, device=dev)
        self.shift = torch.zeros((self.P, 2), dtype=f32, device=dev)
        # per-position tilt tangents and slice thicknesses (tilted propagators), or None
        self.tilt_tan = self.slice_dz = self.tprops = None
        self.set_tilt(getattr(state, 'tilt', None), state.object.thicknesses)

        # work buffers
        S, B, M, NN, H, W = self.S, self.B, self.M, self.NN, self.H, self.W
        self.psi_slots = S if (psi_slots <= 0 or psi_slots >= S) else 3
        self.psi = torch.empty((self.psi_slots, B, M, NN), dtype=c64, device=dev)
        self.chi_a = torch.empty((B, M, NN), dtype=c64, device=dev)
        self.chi_b = torch.empty((B, M, NN), dtype=c64, device=dev)
        self.farfield = torch.empty((B, M, NN), dtype=c64, device=dev)
        # one term per mode and point, the M terms of a point side by side: the forward
        # kernels store |F_m|^2 / n^2 and the readers add the modes in order (no atomics:
        # the model intensity is the same on every run)
        self.intensity = torch.empty((B, NN, self.M), dtype=f32, device=dev)
        self.gpm = torch.empty((S, H, W), dtype=f32, device=dev)
        self.d_probe = torch.zeros((M, NN), dtype=c64, device=dev)
        self.probe_k = torch.empty((M, NN), dtype=c64, device=dev)
        # two object-magnitude planes, alternating by group slot: the clear of group g+1's
        # plane follows group g's backward sweep while group g's probe step still reads its own
        self.gom = torch.empty((2, NN), dtype=f32, device=dev)
        nred = int(_lib.load().phx_red_doubles(S, B, M))
        self.red = torch.zeros((self.G, nred), dtype=torch.float64, device=dev)
        self.errors = torch.zeros((self.G, B), dtype=f32, device=dev)
        self.aux = torch.zeros(2 * B, dtype=f32, device=dev)
        self.finite_flag = torch.full((1,), INT32_MAX, dtype=i32, device=dev)
        self.pos_update = torch.zeros((self.P, 2), dtype=f32, device=dev)
        self.grad_obj = torch.zeros((S, H, W), dtype=c64, device=dev) if gradients else None
        self.grad_probe = torch.zeros((M, NN), dtype=c64, device=dev) if gradients else None
        # LSQ-ML keeps every slice's chi so one launch updates all slices of a group
        self.chi_slices = (torch.empty((S - 1, B, M, NN), dtype=c64, device=dev)
                           if keep_chi and S > 1 else None)
        # planes too large for LDS are transposed through one scratch plane per workgroup
        # the work plane of the probe step in strips (n <= PROBE_STRIPS_MAX_N)
        self.probe_strips = n <= PROBE_STRIPS_MAX_N
        self.probe_tmp = torch.empty((M, NN), dtype=c64, device=dev) if self.probe_strips else None

        # whole-sweep kernels (one launch per sweep) instead of one launch per slice.  A
        # 192 x 192 plane (48 points per lane at 768 threads, 3 waves per SIMD) leaves too
        # few registers for a sweep that also carries its cutout and propagator loads
        # across slices: those kernels spill, and the per-slice kernels (which do not)
        # run a slice step 1.4-2.5x faster, so planes wider than FUSED_MAX_N use them.
        # At SPLIT_N the "whole-sweep" entry points are launch sequences of split passes
        # that fold each slice step's last row pass into the next step's first (2 passes
        # per step instead of 3)
        self.fused = bool(fused) and (n <= FUSED_MAX_N or n == SPLIT_N)
        # a second stream (default priority) for the probe-side kernels of a group step
        self.side_stream = torch.cuda.Stream(device=dev)
        self.concurrent = True
        # group slot whose accumulators the previous launch_lsqml_group call cleared
        self._begun = None
        self.object = obj
        self.probe = probe
        self.descs = [self._make_desc(g) for g in range(self.G)]
        self.desc_refs = [ctypes.byref(d) for d in self.descs]
        # every launch of the group sequences goes through this callable (name, *args);
        # a profiler may wrap it to bracket each kernel with events
        self.call = _lib.call

    def _make_desc(self, g):
        d = _lib.GroupDesc()
        d.n, d.n_modes, d.n_slices, d.obj_h, d.obj_w = self.n, self.M, self.S, self.H, self.W
        d.bmax, d.group_index, d.n_pos = self.B, g, self.P
        d.group_pos = self.group_pos[g].data_ptr()
        d.group_size = self.group_size[g:g + 1].data_ptr()
        d.start = self.start.data_ptr()
        d.shift = self.shift.data_ptr()
        d.ky_s, d.kx_s, d.twiddle = self.ky_s.data_ptr(), self.kx_s.data_ptr(), self.twiddle.data_ptr()
        d.object, d.probe = self.object.data_ptr(), self.probe.data_ptr()
        d.props = None if self.props_s is None else self.props_s.data_ptr()
        d.patterns_s, d.mask_s, d.params = self.patterns_s.data_ptr(), self.mask_s.data_ptr(), self.params.data_ptr()
        d.psi, d.chi_a, d.chi_b = self.psi.data_ptr(), self.chi_a.data_ptr(), self.chi_b.data_ptr()
        d.farfield, d.intensity, d.gpm = self.farfield.data_ptr(), self.intensity.data_ptr(), self.gpm.data_ptr()
        d.d_obj, d.d_probe, d.probe_k = None, self.d_probe.data_ptr(), self.probe_k.data_ptr()
        d.gom = self.gom[g % 2].data_ptr()
        d.red = self.red[g].data_ptr()
        d.errors = self.errors[g].data_ptr()
        d.aux = self.aux.data_ptr()
        d.finite_flag = self.finite_flag.data_ptr()
        d.psi_slots = self.psi_slots
        d.prop_zero_waves = self.prop_zero_waves
        d.grad_obj = None if self.grad_obj is None else self.grad_obj.data_ptr()
        d.grad_probe = None if self.grad_probe is None else self.grad_probe.data_ptr()
        d.chi_slices = None if self.chi_slices is None else self.chi_slices.data_ptr()
        d.tilt = None if self.tilt_tan is None else self.tilt_tan.data_ptr()
        d.slice_dz = None if self.slice_dz is None else self.slice_dz.data_ptr()
        d.tprops = None if self.tprops is None else self.tprops.data_ptr()
        return d

    def set_tilt(self, tilt, thicknesses):
        """Per-position tilts (mrad, (..., 2) in scan order, or None) for the tilted
        propagators of tilt_propagators (simulation.py:193-222): their tangents
        tan(1e-3 tilt) are taken in FP64 here and uploaded as float32 with the slice
        thicknesses.  All-zero tilts give every ramp exactly 1, so they run untilted.
        Call before the descriptors are built, or rebuild them after."""
        tan, dz = tilt_tables(tilt, thicknesses, self.P)
        if tan is None:
            self.tilt_tan = self.slice_dz = self.tprops = None
            return
        self.tilt_tan = self.torch.from_numpy(tan).to(self.device)
        self.slice_dz = self.torch.from_numpy(dz).to(self.device)
        self.tprops = self.torch.empty((self.S - 1, self.B, self.NN), dtype=self.torch.complex64, device=self.device)

    def fill_tilt(self, g, stream=None):
        """The tilted propagators of group slot g's positions into tprops (nothing without
        tilts).  Every forward model of a group starts with it; the backward kernels that
        follow read the same group's copies."""
        if self.tilt_tan is not None:
            self.call("phx_tilt_propagators", self.desc_refs[g], stream_handle() if stream is None else stream)

    # ------------------------------------------------------------------ host tables
    def set_positions(self, obj_sampling, scan):
        """Start indices / subpixel shifts for every position (FP64 host math, uploaded)."""
        scan = numpy.asarray(scan).reshape(-1, 2)
        start, shift = obj_sampling.cutout_indices(scan, (self.n, self.n))
        if numpy.any(start < 0) or numpy.any(start[:, 0] + self.n > self.H) or numpy.any(start[:, 1] + self.n > self.W):
            raise ValueError("probe cutout extends outside the object (the numpy backend asserts start >= 0, "
                             "object.py:318); pad the object")
        self.start.copy_(self.torch.from_numpy(start.astype(numpy.int32)))
        self.shift.copy_(self.torch.from_numpy(shift.astype(numpy.float32)))

    def set_groups(self, groups):
        """Upload the iteration's groups (list of index arrays into the scan shape)."""
        if len(groups) != self.G:
            raise ValueError(f"expected {self.G} groups, got {len(groups)}")
        table = numpy.zeros((self.G, self.B), dtype=numpy.int32)
        sizes = numpy.zeros(self.G, dtype=numpy.int32)
        for g, grp in enumerate(groups):
            flat = numpy.ravel_multi_index(tuple(numpy.asarray(grp)), self.scan_shape)
            if len(flat) > self.B:
                raise ValueError(f"group of {len(flat)} exceeds the {self.B} slots")
            table[g, :len(flat)] = flat
            sizes[g] = len(flat)
        self.group_pos.copy_(self.torch.from_numpy(table))
        self.group_size.copy_(self.torch.from_numpy(sizes))
        self.sizes_host = sizes
        self.table_host = table

    def set_params(self, beta_object=0.0, beta_probe=0.0, reg_object=0.0, reg_probe=0.0, gamma=0.0):
        self.params_host[:5] = (beta_object, beta_probe, reg_object, reg_probe, gamma)
        self.params.copy_(self.torch.from_numpy(self.params_host))

    def reset_finite_flag(self):
        _lib.call("phx_fill_u32", _ptr(self.finite_flag), INT32_MAX, 1, stream_handle())

    def check_state_finite(self, value=0):
        """Record `value` in the finite flag if the object or probe already holds NaN/inf
        (launched before an iteration's first group: the reference's check after that group
        covers the whole state, solvers.py:109)."""
        st = stream_handle()
        for t in (self.object, self.probe):
            self.call("phx_check_finite_f32", _ptr(t), 2 * t.numel(), _ptr(self.finite_flag), int(value), st)

    def finite_group(self):
        v = int(self.finite_flag.item())
        return None if v == INT32_MAX else v

    # ------------------------------------------------------------------ launches
    def launch_dry_run(self, g):
        d, st = self.desc_refs[g], stream_handle()
        c = self.call
        self._begun = None
        c("phx_group_begin", d, st)
        self.fill_tilt(g, st)
        c("phx_probe_spectra", d, st)
        c("phx_shift_probes", d, 0, int(self.S > 1), st)
        for i in range(1, self.S - 1):
            c("phx_forward_slice", d, i, 0, st)
        c("phx_dry_run_stats", d, st)

    def launch_forward(self, g, calc_error, illum=True, modulus=True):
        """Forward model of group slot g up to the exit spectrum and intensity, then (with
        `modulus`) the far-field modulus update into chi_a."""
        d, st = self.desc_refs[g], stream_handle()
        c = self.call
        self._begun = None
        c("phx_group_begin", d, st)
        self.fill_tilt(g, st)
        c("phx_probe_spectra", d, st)
        if self.fused and not illum and self.psi_slots >= self.S:
            c("phx_group_forward", d, st)   # all slices in one launch
            if modulus:
                c("phx_farfield_update", d, int(calc_error), st)
            return
        c("phx_shift_probes", d, int(illum), int(self.S > 1), st)
        for i in range(1, self.S - 1):
            c("phx_forward_slice", d, i, int(illum), st)
        c("phx_farfield", d, int(illum and self.S > 1), int(calc_error), st)

    def _launch_constraints(self, constraints, g, main, side):
        """Group constraints after group g's updates (solvers.py:109-113): object ones on
        the main stream after the object update, probe ones after the probe step (on the
        second stream when it ran there).  The reference checks finiteness before applying
        them, so a NaN they make is reported at the next group (flag g + 1; after the
        last group, the next iteration's state check reports it)."""
        flag = g + 1 if g + 1 < self.G else None
        for con in constraints:
            con.apply(self, g, flag, main if con.target == 'object' else side)

    def launch_lsqml_group(self, g, *, update_object, update_probe, update_position, calc_error, constraints=()):
        """lsqml_run (solvers.py:171-290) for group slot g, then its group constraints.
        Groups are launched in order 0, 1, .. within an iteration (the fused path
        transforms the probe only at g = 0)."""
        d, st = self.desc_refs[g], stream_handle()
        c = self.call
        if self.fused and (self.chi_slices is not None or self.S == 1):
            # forward and backward sweeps in one launch each; every update follows the
            # backward sweep, which reads only pre-update cutouts (solvers.py:246-260).
            # The probe-side kernels (few workgroups each) run on a second stream beside
            # the sweeps and the object update: group_obj_mag needs only the zeroed
            # accumulators and the pre-update object, the probe step only the sweep's
            # probe spectra and alpha_P sums; both streams join before the next group.
            #
            # Capture order decides the hardware queue of every graph node: a node's first
            # captured successor runs on its queue, later ones on other queues.  The order
            # below keeps the chain begin -> forward -> backward -> update -> next forward
            # on one queue (no cross-queue hop between the update and the next forward) and
            # the probe-side kernels on a second one.  Group g+1's accumulators (its own
            # reduction slot and error row, the object-magnitude plane of its
            # parity) are free once group g's backward sweep has run, so they are cleared
            # there; the streams join before each forward sweep, which needs the probe
            # spectra of the previous group's probe step, and after the last group.
            # (an iteration without a probe update runs on the caller's stream alone)
            side = self.side_stream if self.concurrent and update_probe else None
            main = self.torch.cuda.current_stream()
            sh = ctypes.c_void_p(side.cuda_stream) if side is not None else st
            if g == 0 or self._begun != g:
                c("phx_group_begin", d, st)
            if side is not None:
                main.wait_stream(side)
                # group_obj_mag needs the cleared accumulator and the object after the
                # previous group's update, nothing of this group's forward sweep
                side.wait_stream(main)
            if g == 0:
                # later groups find probe_k current: the previous group's probe step
                # writes fft2 of the probe it updates, and without a probe update the
                # probe is the one group 0 transformed
                c("phx_probe_spectra", d, st)
            self.fill_tilt(g, st)
            c("phx_group_forward", d, st)
            if update_probe:
                c("phx_group_obj_mag", d, sh)
            if update_position and self.n == SPLIT_N:
                # split planes: the position step's work plane is the farfield slot, which
                # slice 0's backward step then fills with the probe spectra, so the
                # backward runs per slice with the position step before slice 0
                c("phx_farfield_update", d, int(calc_error), st)
                for i in range(self.S - 1, -1, -1):
                    if i == 0:
                        self._launch_position(d, st)
                    c("phx_lsqml_backward", d, i, int(update_object), int(update_probe), st)
            else:
                c("phx_lsqml_backward_sweep", d, int(update_object), int(update_probe), int(calc_error), st)
                if update_position:
                    self._launch_position(d, st)
            if side is not None:
                side.wait_stream(main)
            self._begun = None
            # with the strip probe step and an object update, the update launch itself
            # clears group g+1's accumulators (none of which it reads)
            fold_begin = FOLD_GROUP_BEGIN and self.probe_strips and update_object and g + 1 < self.G
            if g + 1 < self.G and not fold_begin:
                c("phx_group_begin", self.desc_refs[g + 1], st)
                self._begun = g + 1
            if self.probe_strips:
                # the probe step in strips of small workgroups, which run beside the object
                # update as its workgroups retire: the update starts right away
                if fold_begin:
                    # after the position step, which reads the pre-update slice-0 cutout
                    # (solvers.py:275-288)
                    c("phx_lsqml_object_update_slices_begin", d, 0, self.S, self.desc_refs[g + 1], st)
                    self._begun = g + 1
                elif update_object:
                    c("phx_lsqml_object_update_slices", d, 0, self.S, st)
                if update_probe:
                    c("phx_probe_spectrum_reduce", d, sh)
                    c("phx_lsqml_probe_apply_strips", d, _ptr(self.probe_tmp), sh)
            else:
                if update_probe:
                    c("phx_probe_spectrum_reduce", d, sh)
                    if side is not None and update_object:
                        # the object update starts after the reduction, so the probe step
                        # (four workgroups that each need a whole CU) is dispatched before
                        # the update fills every CU
                        main.wait_stream(side)
                    c("phx_lsqml_probe_apply", d, sh)
                if update_object:
                    c("phx_lsqml_object_update_slices", d, 0, self.S, st)
            self._launch_constraints(constraints, g, st, sh)
            if side is not None and g == self.G - 1:
                main.wait_stream(side)
            return
        # illumination sums are gathered by the object update, not scattered here
        self.launch_forward(g, calc_error, illum=False)
        if update_probe:
            c("phx_group_obj_mag", d, st)
        deferred = self.chi_slices is not None or self.S == 1
        split = self.n == SPLIT_N
        for i in range(self.S - 1, -1, -1):
            if i == 0 and update_position and split:
                # split planes: the position step's work plane is the farfield slot, which
                # slice 0's step then fills with the probe spectra (the position step reads
                # only chi_0, psi_0 and the pre-update slice-0 cutout)
                self._launch_position(d, st)
            c("phx_lsqml_backward", d, i, int(update_object), int(update_probe), st)
            if i == 0 and update_position and not split:
                # the position step uses the group's pre-update slice-0 cutout and the
                # final chi (solvers.py:275-288), so it runs before slice 0 is updated
                self._launch_position(d, st)
            if update_object and not deferred:
                c("phx_lsqml_object_update", d, i, st)
        if update_object and deferred:
            # the backward sweep reads only pre-update cutouts (solvers.py:246-260), so
            # every slice's update can follow it, all in one launch
            c("phx_lsqml_object_update_slices", d, 0, self.S, st)
        if update_probe:
            c("phx_probe_spectrum_reduce", d, st)
            c("phx_lsqml_probe_apply", d, st)
        self._launch_constraints(constraints, g, st, st)

    def _launch_position(self, d, st):
        self.call("phx_position_gradient", d, 0, st)
        self.call("phx_position_gradient", d, 1, st)
        self.call("phx_position_finish", d, _ptr(self.pos_update), st)

    def launch_epie_group(self, g, *, update_object, update_probe, update_position, calc_error, constraints=()):
        """epie_run (solvers.py:415-504) for group slot g, then its group constraints
        (solvers.py:367-371)."""
        d, st = self.desc_refs[g], stream_handle()
        c = self.call
        self.launch_forward(g, calc_error, illum=False)
        for i in range(self.S - 1, -1, -1):
            c("phx_epie_backward", d, i, int(update_object), int(update_probe), st)
            if i == 0 and update_position:
                self._launch_position(d, st)
            if update_object:
                c("phx_epie_object_apply", d, i, st)
        if update_probe:
            c("phx_epie_probe_apply", d, st)
        self._launch_constraints(constraints, g, st, st)

    def scale_probe(self, factor):
        _lib.call("phx_scale_c64", _ptr(self.probe), self.probe.numel(), float(factor), stream_handle())
