/* C ABI of libphaser_mi355x.so: HIP (gfx950) kernels for phaser's reconstruction hot path.
 *
 * Every entry point takes device pointers, plain sizes and a hipStream_t, launches its
 * kernel(s) on that stream and returns 0 on success, PHX_EINVAL (-1000) for a bad
 * argument, PHX_EUNSUPPORTED (-1001) for a size without a kernel, PHX_ETOOLARGE (-1002)
 * for a size beyond a fixed-size device array of the library, or a negative
 * hipError_t.  Nothing allocates, nothing synchronises, no C++ exception crosses the ABI.
 * Complex arrays are interleaved complex64 (float re, float im), C-contiguous.
 *
 * Reference citations are to wdwzyyg/phaser.
 */
#ifndef PHASER_MI355X_H
#define PHASER_MI355X_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------
 * Object cutouts.  Replace the cupy kernels get_cutouts / add_cutouts / set_cutouts
 * (phaser/utils/_cuda_kernels.py:19-107, kernel :163-211) behind ObjectCutout.get/add/set
 * (phaser/utils/object.py:501-557).  start_idxs: (n_cutouts, 2) int64 (y, x) on device;
 * cutouts: (n_cutouts, n_objects, cut_h, cut_w); obj: (n_objects, obj_h, obj_w).
 * Out-of-bounds reads give the default value; out-of-bounds adds/sets are skipped.
 * add_* uses hardware float atomics (summation order is unspecified).
 * --------------------------------------------------------------------------------- */
int phx_get_cutouts_f32(const void* obj, void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        float default_value, hipStream_t stream);
int phx_get_cutouts_c64(const void* obj, void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        float default_re, float default_im, hipStream_t stream);
int phx_add_cutouts_f32(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream);
int phx_add_cutouts_c64(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream);
int phx_set_cutouts_f32(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream);
int phx_set_cutouts_c64(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream);

/* Fill `count` 32-bit words with `value` (a kernel, never a memset node: safe in graphs). */
int phx_fill_u32(void* ptr, uint32_t value, int64_t count, hipStream_t stream);

/* Batched fft2 (inverse=0) / ifft2 (inverse=1) with phaser's convention
 * (phaser/utils/num.py:434-468: centred real space, corner-origin reciprocal space,
 * norm='ortho') over `batch` n x n planes, n in {16, 32, 64, 128, 192, 256}.  `twiddle` holds
 * the n complex64 values exp(-2 pi i j / n).  in != out.  n = 256 runs as two launches (rows
 * into `out`, then columns in place in `out`) and takes batch <= 2^20. */
int phx_fft2_c64(const void* in, void* out, int64_t batch, int64_t n, int inverse, const void* twiddle,
                 hipStream_t stream);

/* Copy n x n planes between natural and scrambled frequency order (see below):
 * unscramble == 0: out[b][py][px] = in[b][f(py)][f(px)]; unscramble != 0: the inverse.
 * is_complex selects complex64 (else float32) elements.  in != out, batch <= 65535. */
int phx_scramble(const void* in, void* out, int64_t batch, int64_t n, int is_complex, int unscramble,
                 hipStream_t stream);

/* check_finite (phaser/utils/num.py:535-541) as a device flag: *flag = min(*flag, value)
 * when any of `count` floats at ptr is NaN or inf.  The engines run it over the object
 * and probe before an iteration's first group, so state that is already non-finite is
 * reported as group 0, where solvers.py:109 raises for it. */
int phx_check_finite_f32(const void* ptr, int64_t count, int32_t* flag, int32_t value, hipStream_t stream);

/* ptr[i] *= scale for `count` complex64 values (presolve probe rescale, solvers.py:60). */
int phx_scale_c64(void* ptr, int64_t count, float scale, hipStream_t stream);

/* ---------------------------------------------------------------------------------
 * Group step of the conventional engines (LSQ-ML: phaser/engines/conventional/
 * solvers.py:131-290; ePIE: :386-504) and the forward/adjoint model of the gradient
 * engine (phaser/engines/gradient/run.py:352-438), split into fused plane kernels.
 *
 * Frequency-domain arrays (propagators, patterns, mask, k-grids, probe spectra) are kept
 * in the FFT kernels' "scrambled" order: along each axis of an n-point transform,
 * position p holds frequency (p / 4) + (n / 4) * bitrev2(p % 4), bitrev2 = {0, 2, 1, 3},
 * for n = 16, 32, 64, 192 (with n / 4 = 48), and 2 (p % 8) + (p / 8) % 2 + 16 (p / 16) for
 * n = 128 (phx_scramble converts).  Real-space arrays are natural order.
 * --------------------------------------------------------------------------------- */
typedef struct phx_group_desc {
  int32_t n;          /* plane side: 16, 32, 64, 128 or 192 */
  int32_t n_modes;    /* probe modes M */
  int32_t n_slices;   /* object slices S */
  int32_t obj_h;      /* object slice height H */
  int32_t obj_w;      /* object slice width W */
  int32_t bmax;       /* group slots launched (>= every group's size) */
  int32_t group_index;/* group number, recorded in finite_flag on NaN/inf */
  int32_t n_pos;      /* total scan positions P */
  const int32_t* group_pos;  /* (bmax) position index of each slot of this group */
  const int32_t* group_size; /* device scalar: slots in use */
  const int32_t* start;      /* (P, 2) cutout start (y, x) */
  const float* shift;        /* (P, 2) subpixel shift (y, x), length units */
  const float* ky_s;         /* (n) ky at scrambled row position */
  const float* kx_s;         /* (n) kx at scrambled column position */
  const void* twiddle;       /* (n) complex64 exp(-2 pi i j / n) */
  void* object;              /* (S, H, W) complex64 */
  void* probe;               /* (M, n, n) complex64 */
  const void* props;         /* (S-1, n, n) complex64, scrambled; may be NULL when S == 1 */
  const float* patterns_s;   /* (P, n*n) float32, scrambled */
  const float* mask_s;       /* (n*n) float32, scrambled */
  const float* params;       /* device: [beta_obj, beta_probe, reg_obj, reg_probe, gamma,
                                          noise_offset, noise_eps, pos_scale] */
  void* psi;                 /* (S, bmax, M, n, n) complex64: stored forward waves */
  void* chi_a;               /* (bmax, M, n, n) complex64 */
  void* chi_b;               /* (bmax, M, n, n) complex64 */
  void* farfield;            /* (bmax, M, n*n) complex64, scrambled exit-wave spectra */
  float* intensity;          /* (bmax, n*n, M) float32 per-mode model intensity |F_m|^2 / n^2,
                                scrambled, the M terms of a point side by side; consumers
                                add them in mode order */
  float* gpm;                /* (S, H, W) float32 group illumination */
  void* d_obj;               /* unused (object updates are gathered per tile); may be NULL */
  void* d_probe;             /* (M, n*n) complex64 probe update spectrum (kept zero) */
  void* probe_k;             /* (M, n*n) complex64 probe spectra, scrambled */
  float* gom;                /* (n*n) float32 group object magnitude */
  double* red;               /* reductions, phx_red_doubles(S, bmax, M) doubles */
  float* errors;             /* (bmax) per-slot detector error of this group */
  float* aux;                /* (2*bmax) float32 scratch (ePIE maxima) */
  int32_t* finite_flag;      /* device int: min group index with NaN/inf (INT32_MAX = none) */
  int32_t psi_slots;         /* planes of psi kept per slot: >= S (or <= 0) stores every slice;
                                3 keeps psi_0 plus two ping-pong slices (recompute mode) */
  int32_t prop_zero_waves;   /* bit w: every column that wave w of a plane workgroup owns in the
                                column phase is zero in every propagator (band limit), so
                                its column passes of a propagation are skipped; 0 = none */
  void* grad_obj;            /* (S, H, W) complex64 object gradient accumulator (gradient engine) */
  void* grad_probe;          /* (M, n, n) complex64 probe gradient (gradient engine) */
  void* chi_slices;          /* (S-1, bmax, M, n, n) complex64 or NULL.  When set, the
                                backward steps keep every slice's chi (chi_{S-1} stays in
                                chi_a) so all object updates of a group can run at once */
  const float* tilt;         /* (P, 2) float32 tangents (tan(1e-3 tilt_y), tan(1e-3 tilt_x)) of
                                each position's tilt in mrad, or NULL for no tilt.  Propagator
                                j of position p is then props[j] exp(2 pi i dz_j (ty ky + tx kx))
                                (tilt_propagators, phaser/engines/common/simulation.py:193-222),
                                and every kernel that propagates reads the group's tilted copies
                                in tprops, which phx_tilt_propagators fills */
  const float* slice_dz;     /* (S-1) float32 slice thicknesses dz_j (length units); read only
                                when tilt is set */
  void* tprops;              /* (S-1, bmax, n*n) complex64, scrambled: the group's tilted
                                propagators (used only when tilt is set) */
} phx_group_desc;

/* Tilted propagators of the group's slots (tilt_propagators, phaser/engines/common/
 * simulation.py:193-222): tprops[j][b] = props[j] exp(2 pi i dz_j (ty ky + tx kx)) with the
 * tangents of slot b's position, the phase taken in half-turns (sincospif).  Launched
 * before a group's forward model (and again before its adjoint if another group was
 * filled in between) when d->tilt is set; PHX_EINVAL when it is not. */
int phx_tilt_propagators(const phx_group_desc* d, hipStream_t stream);

/* Doubles needed by phx_group_desc.red. */
int64_t phx_red_doubles(int32_t n_slices, int32_t bmax, int32_t n_modes);

/* Zero the per-group accumulators (gom, red, errors).  gpm, used only when a
 * forward kernel is called with illum != 0, is the caller's to clear. */
int phx_group_begin(const phx_group_desc* d, hipStream_t stream);

/* probe_k[m] = FFT(probe[m]) (scrambled, unnormalised).  cutout_group's fft2(probes)
 * (phaser/engines/common/simulation.py:255). */
int phx_probe_spectra(const phx_group_desc* d, hipStream_t stream);

/* gom = sum_b |prod_s O_s(cutout b)|^2 (solvers.py:208), summed in a fixed order (no
 * atomics); gom need not be cleared first. */
int phx_group_obj_mag(const phx_group_desc* d, hipStream_t stream);

/* psi[0] = ifft2(fft2(probe) * shift_filter(b)) (simulation.py:241-260, optics.py:151-170);
 * optional illumination sum_m |psi0|^2 scattered into gpm[0] (solvers.py:213-215);
 * norm[b] = sum |psi0|^2; and, when propagate != 0, the slice-0 step
 * psi[1] = ifft2(fft2(psi0 * O_0) * prop_0) (solvers.py:217-220). */
int phx_shift_probes(const phx_group_desc* d, int illum, int propagate, hipStream_t stream);

/* Slice step i (1 <= i <= S-2): psi[i+1] = ifft2(fft2(psi[i] * O_i) * prop_i), with the
 * illumination of slice i when illum != 0 (solvers.py:210-222). */
int phx_forward_slice(const phx_group_desc* d, int slice, int illum, hipStream_t stream);

/* Exit wave (solvers.py:231-240): farfield = FFT(psi[S-1] * O_{S-1}), intensity[b][e][m] =
 * |fft2|^2 / n^2 per mode (summed over m in mode order where read); illumination of slice S-1 when illum != 0.  Then the amplitude/Anscombe
 * modulus constraint (noise_models.py:47-63) chi_a = ifft2(update * Psi) and, when
 * calc_error != 0, errors[b] = sqrt(nansum((I - D)^2)) (solvers.py:237, FP64 sum). */
int phx_farfield(const phx_group_desc* d, int illum, int calc_error, hipStream_t stream);

/* Fused forward model of a group (simulation.py:241-285, solvers.py:202-236): one launch
 * carries each (slot, mode) plane through every slice with the wave in registers:
 * psi[0] = ifft2(fft2(probe) * shift_filter(b)), psi[i+1] = ifft2(fft2(psi[i] * O_i) *
 * prop_i) stored for every i, then farfield = FFT(psi[S-1] * O_{S-1}) and intensity[b][e][m] =
 * |farfield|^2 / n^2 (plain stores, one term per mode), plus the per-slot norm
 * sum_m |psi[0]|^2.  Replaces phx_shift_probes + phx_forward_slice + the first half of
 * phx_farfield.  Needs psi_slots covering every slice. */
int phx_group_forward(const phx_group_desc* d, hipStream_t stream);
/* The far-field modulus update alone (the second half of phx_farfield): chi_a =
 * ifft2((sqrt(D + c)/(sqrt(I + c) + eps) - 1) mask farfield), errors[b] for mode 0 when
 * calc_error (solvers.py:231-240, noise_models.py:47-63). */
int phx_farfield_update(const phx_group_desc* d, int calc_error, hipStream_t stream);
/* Fused LSQ-ML backward sweep (solvers.py:231-271) after phx_group_forward: the modulus
 * update, then for i = S-1 .. 0 with chi_i carried in registers: chi_i stored (chi_a for
 * i = S-1, chi_slices otherwise), the FP64 alpha_O sums when update_object, and
 * chi_{i-1} = ifft2(fft2(chi_i conj(O_i)) conj(prop_{i-1})); at slice 0 with
 * update_probe, each slot's fft2(chi_0 conj(O_0)) conj(filt_b) into farfield and the
 * alpha_P sums.  Object updates (phx_lsqml_object_update_slices) and the probe update
 * follow.  Needs chi_slices when S > 1. */
int phx_lsqml_backward_sweep(const phx_group_desc* d, int update_object, int update_probe, int calc_error,
                             hipStream_t stream);

/* LSQ-ML backward step at slice i (solvers.py:242-271).  Reads chi from chi_a (even
 * steps from the exit) or chi_b; when update_object, accumulates the FP64 alpha_O sums;
 * for i > 0 writes the back-propagated chi to the other buffer; for i == 0 and
 * update_probe, writes each slot's fft2(dP) conj(filt) to farfield and the FP64 alpha_P
 * sums. */
int phx_lsqml_backward(const phx_group_desc* d, int slice, int update_object, int update_probe,
                       hipStream_t stream);
/* LSQ-ML object update of slice i (solvers.py:213-215, 248-257):
 * O_i += beta_O (sum_b alpha_O,b) dO/(gpm + reg) * gpm/(gpm + 1e-16) with
 * dO = sum_{b,m} scatter(chi_i conj(psi_i)) and gpm = sum_{b,m} scatter(|psi_i|^2), both
 * gathered per 64 x 4 object tile in position order (no atomics, deterministic).
 * Without chi_slices it runs after phx_lsqml_backward(i) and before the step of slice
 * i-1 (chi ping-pongs).  With chi_slices it may run after the whole backward sweep: the
 * sweep reads only the pre-update cutouts (solvers.py:246-260), so deferring every slice's
 * update to the end of the group gives the same result. */
int phx_lsqml_object_update(const phx_group_desc* d, int slice, hipStream_t stream);
/* The updates of slices first .. first+count-1 in one launch (needs chi_slices when
 * count > 1). */
int phx_lsqml_object_update_slices(const phx_group_desc* d, int first, int count, hipStream_t stream);
/* phx_lsqml_object_update_slices(d, first, count) and phx_group_begin(next) in one launch:
 * the update of group d also zeroes the accumulators of the group `next` starts from (the
 * fresh per-group arrays of lsqml_run, solvers.py:171-200).  Valid when d's update reads
 * none of next's buffers: next->red must differ from d->red; gom and errors are not read
 * by the update.  Saves the separate clear launch between a group's backward
 * sweep and its object update. */
int phx_lsqml_object_update_slices_begin(const phx_group_desc* d, int first, int count, const phx_group_desc* next,
                                         hipStream_t stream);
/* d_probe[m] = sum_b farfield[b][m] (position order): the probe-update spectrum. */
int phx_probe_spectrum_reduce(const phx_group_desc* d, hipStream_t stream);
/* P += beta_P (sum_b alpha_P) ifft2(d_probe)/(gom + reg) * gom/(gom + 1e-16); d_probe = 0;
 * then probe_k = fft2(P) of the updated probe (as phx_probe_spectra). */
int phx_lsqml_probe_apply(const phx_group_desc* d, hipStream_t stream);
/* The same probe step (same result) in three launches of small workgroups (strips of
 * lines of the plane) that run beside a concurrent object update instead of each
 * needing a whole CU; tmp is an n_modes * n * n complex64 work plane.  n <= 128. */
int phx_lsqml_probe_apply_strips(const phx_group_desc* d, void* tmp, hipStream_t stream);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * probe step of phx_lsqml_probe_apply_strips (same kernels, same values) with its work
 * plane in a fixed device array of the library instead of a `tmp` argument.  It calls no
 * hipMalloc/hipFree.  Limit: n <= 128 and n_modes * n * n <= phx_probe_strip_scratch_elems()
 * (131072 complex64, 1 MiB), else PHX_ETOOLARGE.  Every call on one GPU uses that GPU's
 * single array, so calls that overlap on one GPU (other streams, other threads) race;
 * serialise them.  The array's device address is looked up on the first call on each
 * device (that device must be current). */
int phx_lsqml_probe_apply_strips_static(const phx_group_desc* d, hipStream_t stream);
/* Capacity, in complex64 elements, of phx_lsqml_probe_apply_strips_static's work array. */
int64_t phx_probe_strip_scratch_elems(void);

/* ePIE backward step at slice i (solvers.py:459-485), maxima first (into aux), then the
 * back-propagation of chi through slice i.  At slice 0 with update_probe each plane's
 * probe-update spectrum fft2(P_upd) conj(filt_b) goes to its farfield slot (the far-field
 * step is done with it).  update_object is accepted for the interface and needs nothing
 * here: phx_epie_object_apply gathers the object update afterwards. */
int phx_epie_backward(const phx_group_desc* d, int slice, int update_object, int update_probe,
                      hipStream_t stream);
/* ePIE object update of slice i (solvers.py:466-477): O_i += beta_O/S sum_b (sum_m chi_i
 * conj(psi_i)) / max(sum_m |psi_i|^2)_b over the group's cutouts, gathered per object tile
 * in group order (no atomics).  Call it after phx_epie_backward(slice) and before
 * phx_epie_backward(slice - 1), whose output overwrites chi_i. */
int phx_epie_object_apply(const phx_group_desc* d, int slice, hipStream_t stream);
/* d_probe = sum_b farfield[b] in position order (phx_probe_spectrum_reduce), then
 * P += beta_P ifft2(d_probe / B); d_probe = 0. */
int phx_epie_probe_apply(const phx_group_desc* d, hipStream_t stream);

/* Dry run rescale factors (solvers.py:131-167): rescale[b] = sum(D*mask)/sum|psi[S-1]|^2
 * into errors[b] (by Parseval sum|fft2(psi)|^2 = sum|psi|^2). */
int phx_dry_run_stats(const phx_group_desc* d, hipStream_t stream);

/* Position gradient numerators/denominators (solvers.py:275-288) for direction dir
 * (0 = y, 1 = x), using the final chi (in chi_a when S is odd, chi_b when even).  At
 * n = 256 the farfield slots are its work plane: call it before phx_lsqml_backward of
 * slice 0 (which writes the probe spectra there; the step reads only chi_0, psi_0 and
 * the pre-update slice-0 cutout). */
int phx_position_gradient(const phx_group_desc* d, int dir, hipStream_t stream);

/* ---------------------------------------------------------------------------------
 * Gradient engine (phaser/engines/gradient/run.py:293-438): hand-written adjoint of the
 * multislice model.  Gradients follow torch's convention g = 2 dL/dconj(z) (the reference
 * uses jax's, conj(g), and steps along -conj(g_jax) = -g).
 * --------------------------------------------------------------------------------- */
/* Exit spectrum + intensity, then chi_a = dL/d(exit wave) of the amplitude/Anscombe
 * (poisson = 0, noise_models.py:29-45) or Poisson (poisson = 1, :86-102) loss times
 * params[8]; the group loss (FP64) is added to the group's red.loss slot. */
int phx_farfield_gradient(const phx_group_desc* d, int poisson, hipStream_t stream);
/* Adjoint slice step i (reads the gradient from chi_a/chi_b like phx_lsqml_backward):
 * object gradient into grad_obj[i]; for i > 0 back-propagates; for i == 0 writes each
 * plane's probe-gradient spectrum to its farfield slot (for phx_gradient_probe_finish)
 * and accumulates the position gradients in red.pos_num. */
int phx_gradient_backward(const phx_group_desc* d, int slice, int want_object, int want_probe, int want_positions,
                          hipStream_t stream);
/* The far-field gradient update alone (after phx_group_forward): chi_a = dL/d(exit wave)
 * and the group loss, as the second half of phx_farfield_gradient. */
int phx_farfield_gradient_update(const phx_group_desc* d, int poisson, hipStream_t stream);
/* Fused adjoint sweep (gradient/run.py:293-349 through the hand-written adjoint of the
 * forward model): from the exit-wave gradient in chi_a, every slice in one launch with
 * the gradient carried in registers: with want_object the wave gradient g at slice i is
 * stored as chi_i (chi_slices[i], chi_a at i = S-1) for phx_gradient_object_gather,
 * g <- g conj(O_i), g <- ifft2(fft2(g) conj(prop_{i-1})); at slice 0 each plane's
 * probe-gradient spectrum (into its farfield slot, summed by phx_gradient_probe_finish)
 * and the position gradients (red pos_num).  Needs every psi_i
 * stored (not the recompute mode) and, for the object gradient with S > 1, chi_slices. */
int phx_gradient_sweep(const phx_group_desc* d, int want_object, int want_probe, int want_positions,
                       hipStream_t stream);
/* phx_farfield_gradient_update and phx_gradient_sweep in one launch (the same values):
 * the far-field gradient step and the group loss computed in the sweep's first column
 * pass, the exit-wave gradient transformed in registers (and stored in chi_a with
 * want_object).  For callers that need the loss only after the adjoint pass. */
int phx_gradient_sweep_farfield(const phx_group_desc* d, int poisson, int want_object, int want_probe,
                                int want_positions, hipStream_t stream);
/* phx_group_forward followed by phx_gradient_sweep_farfield (the same values; the loss
 * and the position gradients are FP64 sums with atomics), the loss and gradient of a
 * group (gradient/run.py:293-349).  At 128^2 with more planes (bmax * n_modes) than the
 * device has CUs, the two sweeps run as plane tasks in launches of one workgroup per
 * CU: a launch holds the remaining forward sweeps, then adjoint sweeps of slots whose
 * forward sweeps all ran in an earlier launch (384 planes on 256 CUs: 3 launches for
 * the 4 rounds of the two sweeps).  Otherwise it launches the two sweeps. */
int phx_gradient_group_sweeps(const phx_group_desc* d, int poisson, int want_object, int want_probe,
                              int want_positions, hipStream_t stream);
/* The plane-task schedule phx_gradient_group_sweeps runs (host only, no device call; no
 * counterpart in the reference): for `planes` = bmax * n_modes planes on `cus` CUs,
 * writes 4 ints per launch to plan (first forward plane, forward count, first adjoint
 * plane, adjoint count) and returns the number of launches, or 0 when the two sweeps
 * as separate launches take no more rounds (planes <= cus or planes > 2 cus, or no
 * schedule within max_launches).  PHX_EINVAL for planes not a multiple of modes. */
int phx_plane_task_plan(int planes, int modes, int cus, int* plan, int max_launches);
/* grad_obj[i] = sum_{b,m} scatter(g_i conj(psi_i)) for slices first .. first+count-1
 * from the wave gradients phx_gradient_sweep stored, gathered per object tile in group
 * and mode order (no atomics; deterministic).  Writes every pixel of those slices (zero
 * where no cutout covers it). */
int phx_gradient_object_gather(const phx_group_desc* d, int first, int count, hipStream_t stream);
/* phx_gradient_object_gather(d, first, count) and phx_group_begin(next) in one launch:
 * the gather of group d also zeroes the accumulators of the group `next` starts from
 * (gradient/run.py:293-349 runs the groups of an iteration in turn).  next->red must
 * differ from d->red; the gather reads none of next's buffers. */
int phx_gradient_object_gather_begin(const phx_group_desc* d, int first, int count, const phx_group_desc* next,
                                     hipStream_t stream);
/* d_probe = sum_b farfield[b] in position order (the per-position spectra of the adjoint's
 * slice 0; no atomics, so the result does not depend on scheduling), then
 * grad_probe = ifft2(d_probe) (ortho); d_probe = 0. */
int phx_gradient_probe_finish(const phx_group_desc* d, hipStream_t stream);
/* Gradient-engine rescale factors into errors[b]: sum(D) / sum |psi_{S-1} O_{S-1}|^2
 * (gradient/run.py:408-438: last slice included, pattern mask not applied). */
int phx_dry_run_stats_gradient(const phx_group_desc* d, hipStream_t stream);
/* The same ratio from the model intensity phx_group_forward left (sum over the plane of
 * sum_m |F_m|^2 / n^2 = sum_m |psi_{S-1} O_{S-1}|^2 by Parseval): the dry run as one
 * forward launch and this reduction. */
int phx_dry_run_stats_intensity(const phx_group_desc* d, hipStream_t stream);

/* ---------------------------------------------------------------------------------
 * Constraints and cost regularizers (phaser/engines/common/regularizers.py), applied
 * after each group (group constraints: solvers.py:113, gradient/run.py:344-347), after
 * each iteration (iter constraints: conventional/run.py:93, gradient/run.py:276-279) or in
 * the gradient engine's loss (regularizers: gradient/run.py:399-403).
 * --------------------------------------------------------------------------------- */
/* ClampObjectAmplitude (regularizers.py:20-42): o *= min(|o|, amplitude) / |o| over
 * `count` complex64 values (|o| = 0 gives NaN, as numpy).  With a flag, a non-finite
 * result records flag_value (min). */
int phx_clamp_amplitude_c64(void* obj, int64_t count, float amplitude, int32_t* flag, int32_t flag_value,
                            hipStream_t stream);
/* LimitProbeSupport (regularizers.py:45-64): every probe mode P <- ifft2(fft2(P) mask) with
 * mask (n*n, 0/1, scrambled order) the k-disk of max_angle; probe_k = fft2 of the result. */
int phx_probe_limit_support(const phx_group_desc* d, const float* mask_s, hipStream_t stream);
/* RegularizeLayers (regularizers.py:86-132): per pixel, the complex log of the n_slices
 * slices convolved along the slice axis with taps[0 .. 2 radius] (edge slices repeated,
 * scipy.ndimage.convolve1d mode 'nearest'; FP64 sums), exponentiated and blended:
 * o <- weight exp(.) + one_minus_weight o.  n_slices <= 128. */
int phx_regularize_layers_c64(void* obj, int64_t n_slices, int64_t obj_h, int64_t obj_w, const float* taps,
                              int64_t radius, float weight, float one_minus_weight, hipStream_t stream);
/* Object cost regularizers of the gradient engine (regularizers.py:195-226 obj_l1/obj_l2,
 * :297-316 obj_tikh, :319-338 layers_tv, :341-360 layers_tikh): coefs[5] are their costs
 * times the group's cost scale (B / n_positions), in that order (0 = off).  Adds their
 * loss to *loss (FP64) and, with want_grad, their gradient 2 dL/dconj(O) to grad. */
int phx_cost_regularizers_c64(const void* obj, void* grad, int64_t n_slices, int64_t obj_h, int64_t obj_w,
                              const float* coefs, double* loss, int want_grad, hipStream_t stream);
/* ObjLowPass / ObjGaussian (regularizers.py:135-192): every slice O <- ifft2(F fft2(O)) for
 * a real H x W filter F on the corner-origin frequency grid, given transposed (filt_t[kx][ky]).
 * Object sides are arbitrary: each 1-D DFT is Bluestein's, a circular convolution of
 * length L = 2^p >= 2N - 1 (L <= 4096: sides up to 2048) with host tables per side: chirp (N) exp(-i pi k^2/N),
 * bhat (L) FFT_L(conj chirp)/L, tw (L/2) exp(-2 pi i j/L).  work: n_slices*obj_h*obj_w
 * complex64 (the transposed object). */
int phx_object_filter_c64(void* obj, void* work, int64_t n_slices, int64_t obj_h, int64_t obj_w,
                          const float* filt_t, const void* chirp_h, const void* bhat_h, const void* tw_h,
                          int64_t l_h, const void* chirp_w, const void* bhat_w, const void* tw_w, int64_t l_w,
                          hipStream_t stream);

/* Optimizer steps on complex64 parameters (gradient/solvers.py:88-138 over optax): the
 * direction is d = -grad / scale; finite_flag (may be NULL) records flag_value on NaN/inf.
 * SGD: t = d + momentum t; u = nesterov ? d + momentum t : t (or u = d without momentum);
 *      param += lr u. */
int phx_sgd_step(void* param, const void* grad, void* trace, int64_t n, float scale, float lr, float momentum,
                 int nesterov, int use_momentum, int32_t* finite_flag, int32_t flag_value, hipStream_t stream);
/* Adam (optax.scale_by_adam): mu, nu updated in place; count is the step number after
 * increment (1 on the first step); param += lr mu_hat / (sqrt(nu_hat + eps_root) + eps). */
int phx_adam_step(void* param, const void* grad, void* mu, float* nu, int64_t n, float scale, float lr, float b1,
                  float b2, float eps, float eps_root, int32_t count, int nesterov, int32_t* finite_flag,
                  int32_t flag_value, hipStream_t stream);
/* acc += sum |grad / scale|^2 (FP64). */
int phx_sqnorm_accumulate(const void* grad, int64_t n, float scale, double* acc, hipStream_t stream);
/* Polyak SGD (scale_by_learning_rate(scaling) + optax.scale_by_polyak): step =
 * min((loss - f_min) / (scaling^2 sqnorm + eps), max_lr); param += step scaling d. */
int phx_polyak_step(void* param, const void* grad, int64_t n, float scale, float scaling, float max_lr, float f_min,
                    float eps, const double* sqnorm, const double* loss, int32_t* finite_flag, int32_t flag_value,
                    hipStream_t stream);

/* Model intensity sum_m |fft2(psi[S-1] * O_{S-1})|^2 of each slot, written in natural
 * (corner-origin) order to out[group_pos[b]] ((P, n*n) float32): the forward model used to
 * simulate diffraction data (no reference counterpart; the reference's gradient engine
 * evaluates the same model inside its loss, gradient/run.py:385-393). */
int phx_exit_intensity(const phx_group_desc* d, float* out, hipStream_t stream);

/* pos_update[(group_pos[b], dir)] = num[b][dir] / den[dir] for the group's slots, after
 * both phx_position_gradient calls; pos_update is (P, 2) float32. */
int phx_position_finish(const phx_group_desc* d, void* pos_update, hipStream_t stream);

/* ---------------------------------------------------------------------------------
 * Learned initializer (no counterpart in the reference: an MLP that maps binned
 * diffraction patterns to object-patch phase and amplitude, trained on the engines' own
 * reconstructions; its guesses are stitched into the starting object).
 * BF16 operands (raw 16-bit bfloat16), FP32 accumulation.
 * --------------------------------------------------------------------------------- */
/* C[M, N] (FP32 and/or BF16 Cb, leading dimension ldc) = op(A) op(B), op(A) = A (M x K,
 * row-major, lda) or, with trans_a, the transpose of a K x M array; op(B) = B (K x N) or,
 * with trans_b, the transpose of an N x K array.  Epilogue in order: + C when
 * accumulate, + bias[n], max(., 0) when relu, zeroed where mask (BF16, ldm) is <= 0
 * (the ReLU derivative of a stored activation).  M, N multiples of 128; K of 32; lda, ldb
 * multiples of 8. */
int phx_gemm_bf16(const void* A, int64_t lda, int trans_a, const void* B, int64_t ldb, int trans_b, int64_t M,
                  int64_t N, int64_t K, float* C, void* Cb, int64_t ldc, const float* bias, int relu, const void* mask,
                  int64_t ldm, int accumulate, hipStream_t stream);
/* The NT form of phx_gemm_bf16 (A M x K, B N x K, both k-contiguous) with an optional
 * transposed BF16 copy of the result, Ct[n * ldct + m]: the double-buffered kernel every
 * GEMM of the network's training step and inference runs.  M and N multiples of 128, K of
 * 64; phx_gemm_bf16 routes its NT form with K a multiple of 64 here too. */
int phx_gemm_bf16_nt(const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K, float* C,
                     void* Cb, int64_t ldc, void* Ct, int64_t ldct, const float* bias, int relu, const void* mask,
                     int64_t ldm, int accumulate, hipStream_t stream);
/* out[c * ldo + r] = in[r * ldi + c] for a rows x cols BF16 array (in != out). */
int phx_transpose_bf16(const void* in, int64_t ldi, int64_t rows, int64_t cols, void* out, int64_t ldo,
                       hipStream_t stream);
/* out[c] = sum_r in[r * ld + c], rows summed in order (FP32). */
int phx_colsum_f32(const float* in, int64_t rows, int64_t cols, int64_t ld, float* out, hipStream_t stream);
/* Adam on FP32 master weights w (m, v moments; c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)); the BF16
 * copy wb (may be NULL) is refreshed.  g += weight_decay * w first. */
int phx_adam_f32_bf16(float* w, void* wb, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                      float eps, float c1, float c2, float weight_decay, hipStream_t stream);
int phx_f32_to_bf16(const float* in, void* out, int64_t n, hipStream_t stream);
/* d = scale (pred - target) (FP32 d and/or BF16 db may be NULL); *loss += sum (pred -
 * target)^2 in FP64 when loss != NULL. */
int phx_mse_grad(const float* pred, const float* target, int64_t n, float scale, float* d, void* db, double* loss,
                 hipStream_t stream);
/* phx_mse_grad over a rows x cols array, also writing the BF16 transpose dbt (cols x rows;
 * may be NULL, as may d; dbt needs db). */
int phx_mse_grad_t(const float* pred, const float* target, int64_t rows, int64_t cols, float scale, float* d, void* db,
                   void* dbt, double* loss, hipStream_t stream);
/* Network input: out[r] = log(1 + binned(pattern[index[r]]) * inv_scale), the sums of
 * bin x bin pixel blocks, (n/bin)^2 BF16 values per row (n a multiple of bin); patterns
 * (P, n, n) FP32 corner-origin; index NULL means r. */
int phx_pattern_features(const float* patterns, const int32_t* index, int64_t count, int64_t n, int64_t bin,
                         float inv_scale, void* out, hipStream_t stream);
/* Training targets: out[r] = [phase - phase0, amplitude] of prod_s O_s over the ps x ps
 * window starting at starts[r] = (y0, x0) (2 ps^2 FP32 per row). */
int phx_object_patches(const void* obj, int64_t S, int64_t H, int64_t W, const int32_t* starts, int64_t count,
                       int64_t ps, float phase0, float* out, hipStream_t stream);
/* Stitch predicted rows [phase, amplitude] (row stride ldp) at their windows: per-pixel
 * averages (acc_* zeroed by the caller), then obj[s] = a^(1/S) exp(i phi / S) for every
 * slice, 1 where no window covers the pixel. */
int phx_stitch_patches(const float* pred, int64_t ldp, const int32_t* starts, int64_t count, int64_t ps, int64_t S,
                       int64_t H, int64_t W, float* acc_phase, float* acc_amp, float* acc_n, void* obj,
                       hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
