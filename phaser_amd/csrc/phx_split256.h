// Group kernels for 256 x 256 planes (included by phx_engine.hip inside its anonymous
// namespace, after the group helpers it uses: Cut, obj_at, shift_filter, prop_of,
// psi_off, chi_in_buf/chi_out_buf, model_intensity, farfield_factor, red_layout).
//
// A 256 x 256 complex64 plane (512 KiB) fits neither a CU's LDS nor one workgroup's
// registers, so every 2-D transform of a group step is split over launches: a row pass
// (workgroups of 16 whole rows, 256 threads) and a column pass (workgroups of 32 whole
// columns, 512 threads), the plane going through HBM between them.  Element-wise work is
// applied where a pass loads or stores a point (the cutout and transmission before a
// forward row transform, the propagator between a column pass's forward and inverse
// transforms, the modulus constraint before an inverse column transform), so a slice step
// of the forward model is three passes over its plane (rows, columns, rows) and each pass
// reads and writes the plane once.  The wave in flight between passes lives in the
// buffer that receives the step's result (psi[i+1], chi_{i-1}, farfield), so no scratch
// plane is needed.
//
// Layout: every plane in natural order, spectra included (frequency index k at k, the
// host's ky/kx tables and the patterns in that order).  Transforms are raw (unnormalised)
// DFTs, W = exp(-2 pi i / 256) forward; the 1/n^2 of an inverse 2-D transform is applied
// where its row pass stores the result, as the single-workgroup plane kernels do.
//
// One 256-point line is held by 16 threads: thread t of a line holds x[16 a + t] in
// register a, and the transform is the four-step 16 x 16 split
//   X[k1 + 16 k2] = sum_t W_16^(t k2) W_256^(t k1) sum_a x[16 a + t] W_16^(a k1):
// an in-register DFT_16, the twiddle W_256^(t k1), an exchange through LDS after which
// thread t holds the 16 values of k1 = t, and a second in-register DFT_16 that leaves
// X[t + 16 k2] in register k2.  The output index of register k equals the input index
// of register a = k, so a forward transform, a point-wise product and an inverse
// transform chain in registers.

constexpr int S256 = 256;
constexpr int S256_NN = S256 * S256;
constexpr int S256_LD = 17;                    // padded row of a line's 16 x 16 exchange block
constexpr int S256_ROW_LINES = 16;             // rows per row-pass workgroup (256 threads)
constexpr int S256_COL_LINES = 32;             // columns per column-pass workgroup (512 threads)
constexpr int S256_ROW_TILES = S256 / S256_ROW_LINES;
constexpr int S256_COL_TILES = S256 / S256_COL_LINES;
constexpr float S256_INV_NN = 1.0f / (float)S256_NN;

// Exchange blocks and the twiddle table of one pass workgroup.  In the column pass the
// 32 lines of a load are adjacent lanes, so the blocks sit at an odd stride.
template <bool COL>
struct S256Lds {
  static constexpr int LINES = COL ? S256_COL_LINES : S256_ROW_LINES;
  static constexpr int BLK = 16 * S256_LD + (COL ? 1 : 0);
  float2 ex[LINES * BLK];
  float2 tw[S256];
  double red[COL ? 8 : 4];  // one FP64 partial per wave for block sums
};

// Pass geometry: line l of the workgroup, thread t within the line, the line's index in
// the plane, and the plane element held in register a.
template <bool COL>
struct S256Pass {
  int l, t, line;
  __device__ __forceinline__ S256Pass(int tile) {
    const int tid = threadIdx.x;
    l = COL ? (tid % S256_COL_LINES) : (tid / 16);
    t = COL ? (tid / S256_COL_LINES) : (tid % 16);
    line = (COL ? S256_COL_LINES : S256_ROW_LINES) * tile + l;
  }
  // (row, column) of register a
  __device__ __forceinline__ int y(int a) const { return COL ? 16 * a + t : line; }
  __device__ __forceinline__ int x(int a) const { return COL ? line : 16 * a + t; }
  __device__ __forceinline__ int e(int a) const { return y(a) * S256 + x(a); }
};

template <bool COL>
__device__ __forceinline__ void s256_load_twiddles(S256Lds<COL>& s, const phx_group_desc& d) {
  constexpr int T = 16 * S256Lds<COL>::LINES;
  for (int k = threadIdx.x; k < S256; k += T) s.tw[k] = ((const float2*)d.twiddle)[k];
  __syncthreads();
}

// One 256-point transform of every line of the workgroup (see the header comment).  All
// threads of the block call it (it synchronises the block).
template <bool COL, bool INV>
__device__ __forceinline__ void s256_line(float2 (&x)[16], S256Lds<COL>& s, int l, int t) {
  dft<16, INV>(x);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) {
    const float2 w = s.tw[(t * k1) & 255];
    x[k1] = INV ? cmulc(x[k1], w) : cmul(x[k1], w);
  }
  float2* blk = s.ex + l * S256Lds<COL>::BLK;
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) blk[k1 * S256_LD + t] = x[k1];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = blk[t * S256_LD + j];
  __syncthreads();  // the block is free for the next transform
  dft<16, INV>(x);
}

// FP64 sum over the block of a per-thread FP32 partial (all threads call it).
template <bool COL>
__device__ __forceinline__ double s256_block_sum(float v, S256Lds<COL>& s) {
  return block_sum_f64((double)v, s.red);
}

// ----- operations of the passes -----
enum S256Op {
  // row passes
  R_PROBE_FWD = 0,   // probe[m] -> forward rows -> probe_k[m]                        (per mode)
  R_SHIFT_FINISH,    // psi_0 semi -> inverse rows / n^2 -> psi_0, norm; [* O_0 -> fwd rows -> psi_1]
  R_SLICE_START,     // [psi_i semi -> inverse rows / n^2 -> psi_i;] psi_i * O_i -> forward rows -> psi_{i+1}
  R_WAVE_FINISH,     // psi_{i+1} semi -> inverse rows / n^2 -> psi_{i+1}
  R_EXIT,            // psi_{S-1} * O_{S-1} -> forward rows -> farfield
  R_CHI_FINISH,      // chi_a semi -> inverse rows / n^2 -> chi_a
  R_BACK,            // [chi_i semi -> inverse rows / n^2 -> chi_i;] step-size sums, dP = chi conj(O_i)
                     //   -> forward rows -> chi_{i-1} or farfield
  R_BACK_FINISH,     // chi_{i-1} semi -> inverse rows / n^2 -> chi_{i-1}
  R_PROBE_UPDATE,    // d_probe semi -> inverse rows, P += update; P -> forward rows -> d_probe (per mode)
  R_EPIE_BACK,       // ePIE: chi_i conj(O_i) / max|O|^2 -> forward rows -> chi_{i-1} or farfield
  R_EPIE_PROBE,      // ePIE: d_probe semi -> inverse rows, P += beta_P v / (n^2 B); d_probe = 0 (per mode)
  R_GRAD_PROBE,      // gradient engine: d_probe semi -> inverse rows / n^2 -> grad_probe; d_probe = 0 (per mode)
  R_POS_START,       // LSQ-ML position step: psi_0 -> forward rows -> farfield (work plane)
  R_POS_FINISH,      // inverse rows of d psi_0 / d pos, times O_0, dotted with chi_0 -> position sums
  // column passes
  C_PROBE_FWD,       // probe_k[m] forward columns in place                          (per mode)
  C_SHIFT,           // probe_k[m] filt_b -> inverse columns -> psi_0
  C_PROPAGATE,       // forward columns, * p_i (or conj(p_{i-1}) backward), inverse columns, in place
  C_EXIT,            // forward columns -> farfield, this mode's |F|^2 / n^2 intensity terms
  C_MODULUS,         // modulus update of farfield -> inverse columns -> chi_a
  C_BACK_PROBE,      // forward columns of dP, * conj(filt_b) -> farfield (per-position probe spectrum)
  C_PROBE_INV,       // d_probe[m] inverse columns in place                          (per mode)
  C_PROBE_SPEC,      // d_probe[m] forward columns -> probe_k[m]; d_probe = 0        (per mode)
  C_GRAD_AMPLITUDE,  // gradient engine: amplitude-loss gradient of farfield -> inverse columns -> chi_a, loss
  C_GRAD_POISSON,    // the same for the Poisson loss
  C_GRAD_PROBE,      // gradient engine, slice 0: forward columns -> probe-gradient spectrum, position gradients
  C_POS,             // LSQ-ML position step: forward columns, * (-2 pi i k), inverse columns, in place
};

// Row pass.  Grid (S256_ROW_TILES * M, bmax), or (S256_ROW_TILES * M, 1) for the
// per-mode operations; 256 threads.  `i` is the slice; flags: bit 0 propagate
// (R_SHIFT_FINISH) / object sums (R_BACK), bit 1 probe sums and spectrum (R_BACK), bit 2
// the slice-0 spectrum without sums (R_BACK), bit 3 finish first (R_SLICE_START, R_BACK:
// the slice's wave or chi arrives from a column pass and its inverse row transform is
// completed, stored, and used in the same pass).
template <int OP>
__global__ void __launch_bounds__(256) s256_rows_k(phx_group_desc d, int i, int flags) {
  __shared__ S256Lds<false> s;
  const int tile = blockIdx.x % S256_ROW_TILES, m = blockIdx.x / S256_ROW_TILES, b = blockIdx.y;
  constexpr bool PER_MODE = OP == R_PROBE_FWD || OP == R_PROBE_UPDATE || OP == R_EPIE_PROBE || OP == R_GRAD_PROBE;
  if (!PER_MODE && !slot_active(d, b)) return;
  const S256Pass<false> g(tile);
  s256_load_twiddles<false>(s, d);
  const int M = d.n_modes, bmax = d.bmax;
  const size_t plane = ((size_t)b * M + m) * S256_NN;
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  float2 x[16];

  if constexpr (OP == R_PROBE_FWD) {
    const float2* P = (const float2*)d.probe + (size_t)m * S256_NN;
    float2* Pk = (float2*)d.probe_k + (size_t)m * S256_NN;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = P[g.e(a)];
    s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) Pk[g.e(a)] = x[a];
  } else if constexpr (OP == R_SHIFT_FINISH || OP == R_WAVE_FINISH || OP == R_CHI_FINISH || OP == R_BACK_FINISH) {
    // the inverse row transform that completes an inverse 2-D transform, stored / n^2
    float2* w;
    if (OP == R_SHIFT_FINISH) w = (float2*)d.psi + psi_off(d, 0, b, m);
    else if (OP == R_WAVE_FINISH) w = (float2*)d.psi + psi_off(d, i + 1, b, m);
    else if (OP == R_CHI_FINISH) w = (float2*)d.chi_a + plane;
    else w = chi_out_buf(d, i) + plane;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = w[g.e(a)];
    s256_line<false, true>(x, s, g.l, g.t);
    float nrm = 0.f;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      x[a] = cscale(x[a], S256_INV_NN);
      w[g.e(a)] = x[a];
      if (OP == R_SHIFT_FINISH) nrm += abs2f(x[a]);
    }
    if constexpr (OP == R_SHIFT_FINISH) {
      // norm[b] = sum_m |psi_0|^2 (FP64), then the slice-0 step's first pass
      const double tot = s256_block_sum<false>(nrm, s);
      if (threadIdx.x == 0) atomic_add_f64(&d.red[red_layout(d.n_slices, bmax, M).norm + b], tot);
      if (flags & 1) {
        const Cut c = cut_of(d, d.group_pos[b]);
        const float2* obj0 = (const float2*)d.object;
        float2* psi1 = (float2*)d.psi + psi_off(d, 1, b, m);
#pragma unroll
        for (int a = 0; a < 16; ++a) x[a] = cmul(x[a], obj_at(obj0, c, g.y(a), g.x(a)));
        s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
        for (int a = 0; a < 16; ++a) psi1[g.e(a)] = x[a];
      }
    }
  } else if constexpr (OP == R_SLICE_START || OP == R_EXIT) {
    // the last slice's product goes to the farfield slot (the exit wave's first pass)
    const int si = OP == R_EXIT ? d.n_slices - 1 : i;
    const Cut c = cut_of(d, d.group_pos[b]);
    float2* psi_i = (float2*)d.psi + psi_off(d, si, b, m);
    const float2* obj_i = (const float2*)d.object + si * HW;
    float2* out = si == d.n_slices - 1 ? (float2*)d.farfield + plane : (float2*)d.psi + psi_off(d, si + 1, b, m);
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = psi_i[g.e(a)];
    if (flags & 8) {
      s256_line<false, true>(x, s, g.l, g.t);
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        x[a] = cscale(x[a], S256_INV_NN);
        psi_i[g.e(a)] = x[a];
      }
    }
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = cmul(x[a], obj_at(obj_i, c, g.y(a), g.x(a)));
    s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) out[g.e(a)] = x[a];
  } else if constexpr (OP == R_BACK) {
    // LSQ-ML backward step at slice i (solvers.py:242-271), its first pass: the step-size
    // sums of the (pre-update) cutouts and dP = chi conj(O_i), transformed along rows into
    // chi_{i-1}'s buffer (i > 0) or the farfield slot (i == 0 with a probe update, or with
    // flags bit 2: the gradient engine's adjoint, which takes no step-size sums)
    const bool upd_obj = flags & 1, probe = flags & 2, spectrum = flags & 4;
    const Cut c = cut_of(d, d.group_pos[b]);
    float2* chi = chi_in_buf(d, i) + plane;
    const float2* psi_i = (const float2*)d.psi + psi_off(d, i, b, m);
    const float2* obj_i = (const float2*)d.object + i * HW;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = chi[g.e(a)];
    if (flags & 8) {
      s256_line<false, true>(x, s, g.l, g.t);
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        x[a] = cscale(x[a], S256_INV_NN);
        chi[g.e(a)] = x[a];
      }
    }
    float numO = 0.f, denO = 0.f, numP = 0.f, denP = 0.f;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const float2 cv = x[a];
      const float2 o = obj_at(obj_i, c, g.y(a), g.x(a));
      if (upd_obj) {
        const float2 ps = psi_i[g.e(a)];
        const float2 tt = cmul(cmulc(cv, ps), ps);  // (chi conj(psi)) psi
        numO += cv.x * tt.x + cv.y * tt.y;
        denO += abs2f(tt);
      }
      const float2 dP = cmulc(cv, o);
      if (probe) {
        const float2 t2 = cmul(dP, o);
        numP += cv.x * t2.x + cv.y * t2.y;
        denP += abs2f(t2);
      }
      x[a] = dP;
    }
    const RedLayout L = red_layout(d.n_slices, bmax, M);
    if (upd_obj) {
      const double ta = s256_block_sum<false>(numO, s), tb = s256_block_sum<false>(denO, s);
      if (threadIdx.x == 0) {
        atomic_add_f64(&d.red[L.num_obj + i * bmax + b], ta);
        atomic_add_f64(&d.red[L.den_obj + i], tb);
      }
    }
    if (probe) {
      const double ta = s256_block_sum<false>(numP, s), tb = s256_block_sum<false>(denP, s);
      if (threadIdx.x == 0) {
        atomic_add_f64(&d.red[L.numP + b * M + m], ta);
        atomic_add_f64(&d.red[L.denP], tb);
      }
    }
    if (i > 0 || probe || spectrum) {
      float2* out = i > 0 ? chi_out_buf(d, i) + plane : (float2*)d.farfield + plane;
      s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
      for (int a = 0; a < 16; ++a) out[g.e(a)] = x[a];
    }
  } else if constexpr (OP == R_PROBE_UPDATE) {
    // LSQ-ML probe step (solvers.py:262-270), its row pass: P += beta_P alpha_P
    // avg/(gom + reg) * gom/(gom + 1e-16) with avg = ifft2(d_probe), then the forward
    // row transform of the new probe into d_probe (the column pass finishes probe_k)
    __shared__ float sA;
    if (threadIdx.x < 64) {
      const RedLayout L = red_layout(d.n_slices, bmax, M);
      const int gs = *d.group_size;
      double sacc = 0.0;
      for (int bb = threadIdx.x; bb < gs; bb += 64) sacc += d.red[L.numP + bb * M + m];
      sacc = wave_sum_f64(sacc);
      if (threadIdx.x == 0) {
        const double den = d.red[L.denP] + (double)d.params[P_GAMMA];
        sA = (float)((double)d.params[P_BETA_PROBE] * (sacc / den));
      }
    }
    float2* dPk = (float2*)d.d_probe + (size_t)m * S256_NN;
    float2* P = (float2*)d.probe + (size_t)m * S256_NN;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = dPk[g.e(a)];
    s256_line<false, true>(x, s, g.l, g.t);  // its barriers publish sA
    const float A = sA, reg = d.params[P_REG_PROBE];
    bool bad = false;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int e = g.e(a);
      const float gm = d.gom[e];
      const float2 avg = cscale(cscale(x[a], S256_INV_NN), 1.0f / (gm + reg));
      const float2 u = cscale(cscale(cscale(avg, A), gm), 1.0f / (gm + 1e-16f));
      const float2 o = cadd(P[e], u);
      P[e] = o;
      bad |= !(isfinite(o.x) && isfinite(o.y));
      x[a] = o;
    }
    if (bad) atomicMin(d.finite_flag, d.group_index);
    s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) dPk[g.e(a)] = x[a];
  } else if constexpr (OP == R_EPIE_BACK) {
    // ePIE backward step at slice i (solvers.py:459-485), its first pass: the probe-side
    // update chi conj(O_i) / max|O_i|^2 of the (pre-update) cutout, transformed along rows
    // into chi_{i-1}'s buffer (i > 0) or the farfield slot (i == 0, probe update)
    const Cut c = cut_of(d, d.group_pos[b]);
    const float2* chi = chi_in_buf(d, i) + plane;
    const float2* obj_i = (const float2*)d.object + i * HW;
    const float inv_maxo = 1.0f / d.aux[b];
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = cscale(cmulc(chi[g.e(a)], obj_at(obj_i, c, g.y(a), g.x(a))), inv_maxo);
    float2* out = i > 0 ? chi_out_buf(d, i) + plane : (float2*)d.farfield + plane;
    s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) out[g.e(a)] = x[a];
  } else if constexpr (OP == R_EPIE_PROBE) {
    // ePIE probe step (solvers.py:481-484): P += beta_P ifft2(mean_b spectra)
    const float sc = d.params[P_BETA_PROBE] * S256_INV_NN / (float)(*d.group_size);
    float2* dPk = (float2*)d.d_probe + (size_t)m * S256_NN;
    float2* P = (float2*)d.probe + (size_t)m * S256_NN;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      x[a] = dPk[g.e(a)];
      dPk[g.e(a)] = make_float2(0.f, 0.f);  // d_probe is kept zero
    }
    s256_line<false, true>(x, s, g.l, g.t);
    bool bad = false;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const float2 o = cadd(P[g.e(a)], cscale(x[a], sc));
      P[g.e(a)] = o;
      bad |= !(isfinite(o.x) && isfinite(o.y));
    }
    if (bad) atomicMin(d.finite_flag, d.group_index);
  } else if constexpr (OP == R_POS_START) {
    const float2* psi0 = (const float2*)d.psi + psi_off(d, 0, b, m);
    float2* w = (float2*)d.farfield + plane;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = psi0[g.e(a)];
    s256_line<false, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) w[g.e(a)] = x[a];
  } else if constexpr (OP == R_POS_FINISH) {
    // LSQ-ML position step (solvers.py:275-288) in direction flags (0: y, 1: x):
    // prod = ifft2(fft2(psi_0) (-2 pi i k)) / sum_m |psi_0|^2 * O_0 (pre-update cutout),
    // num[b] += Re sum conj(chi_0) prod, den += sum |prod|^2 (FP64)
    const int dir = flags & 1;
    const RedLayout L = red_layout(d.n_slices, bmax, M);
    const Cut c = cut_of(d, d.group_pos[b]);
    const float2* w = (const float2*)d.farfield + plane;
    const float2* chi = chi_in_buf(d, 0) + plane;
    const float2* obj0 = (const float2*)d.object;
    const float scale = (float)(1.0 / ((double)S256_NN * d.red[L.norm + b]));
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = w[g.e(a)];
    s256_line<false, true>(x, s, g.l, g.t);
    float num = 0.f, den = 0.f;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const float2 prod = cmul(cscale(x[a], scale), obj_at(obj0, c, g.y(a), g.x(a)));
      const float2 cv = chi[g.e(a)];
      num += cv.x * prod.x + cv.y * prod.y;
      den += abs2f(prod);
    }
    const double ta = s256_block_sum<false>(num, s), tb = s256_block_sum<false>(den, s);
    if (threadIdx.x == 0) {
      atomic_add_f64(&d.red[L.pos_num + 2 * b + dir], ta);
      atomic_add_f64(&d.red[L.pos_den + dir], tb);
    }
  } else if constexpr (OP == R_GRAD_PROBE) {
    // grad_probe[m] = ifft2(sum_b spectra) (the probe gradient), d_probe kept zero
    float2* dPk = (float2*)d.d_probe + (size_t)m * S256_NN;
    float2* gP = (float2*)d.grad_probe + (size_t)m * S256_NN;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      x[a] = dPk[g.e(a)];
      dPk[g.e(a)] = make_float2(0.f, 0.f);
    }
    s256_line<false, true>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) gP[g.e(a)] = cscale(x[a], S256_INV_NN);
  }
}

// Column pass.  Grid (S256_COL_TILES * M, bmax), or (S256_COL_TILES * M, 1) for the
// per-mode operations; 512 threads.  C_PROPAGATE: `i` is the propagator index and flags
// bit 0 selects conj(p) (back-propagation); the plane is psi[i+1] (forward) or
// chi_out_buf(i + 1) (backward).  C_EXIT / C_MODULUS: flags bit 0 marks the last slice.
template <int OP>
__global__ void __launch_bounds__(512) s256_cols_k(phx_group_desc d, int i, int flags) {
  __shared__ S256Lds<true> s;
  const int tile = blockIdx.x % S256_COL_TILES, m = blockIdx.x / S256_COL_TILES, b = blockIdx.y;
  constexpr bool PER_MODE = OP == C_PROBE_FWD || OP == C_PROBE_INV || OP == C_PROBE_SPEC;
  if (!PER_MODE && !slot_active(d, b)) return;
  const S256Pass<true> g(tile);
  s256_load_twiddles<true>(s, d);
  const int M = d.n_modes;
  const size_t plane = ((size_t)b * M + m) * S256_NN;
  float2 x[16];

  if constexpr (OP == C_PROBE_FWD || OP == C_PROBE_INV) {
    float2* w = OP == C_PROBE_FWD ? (float2*)d.probe_k + (size_t)m * S256_NN : (float2*)d.d_probe + (size_t)m * S256_NN;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = w[g.e(a)];
    s256_line<true, OP == C_PROBE_INV>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) w[g.e(a)] = x[a];
  } else if constexpr (OP == C_PROBE_SPEC) {
    float2* w = (float2*)d.d_probe + (size_t)m * S256_NN;
    float2* Pk = (float2*)d.probe_k + (size_t)m * S256_NN;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = w[g.e(a)];
#pragma unroll
    for (int a = 0; a < 16; ++a) w[g.e(a)] = make_float2(0.f, 0.f);  // d_probe is kept zero
    s256_line<true, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) Pk[g.e(a)] = x[a];
  } else if constexpr (OP == C_SHIFT) {
    // psi_0 = ifft2(fft2(P_m) filt_b) (simulation.py:241-260), its inverse column pass
    const int p = d.group_pos[b];
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    const float2* Pk = (const float2*)d.probe_k + (size_t)m * S256_NN;
    float2* psi0 = (float2*)d.psi + psi_off(d, 0, b, m);
#pragma unroll
    for (int a = 0; a < 16; ++a)
      x[a] = cmul(Pk[g.e(a)], shift_filter(sy, sx, d.ky_s[g.y(a)], d.kx_s[g.x(a)]));
    s256_line<true, true>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) psi0[g.e(a)] = x[a];
  } else if constexpr (OP == C_PROPAGATE) {
    const bool back = flags & 1;
    float2* w = back ? chi_out_buf(d, i + 1) + plane : (float2*)d.psi + psi_off(d, i + 1, b, m);
    const float2* prop = prop_of(d, i, b);
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = w[g.e(a)];
    s256_line<true, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const float2 pv = prop[g.e(a)];
      x[a] = back ? cmulc(x[a], pv) : cmul(x[a], pv);
    }
    s256_line<true, true>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) w[g.e(a)] = x[a];
  } else if constexpr (OP == C_EXIT) {
    float2* F = (float2*)d.farfield + plane;
    float* I = intensity_term(d, b, 0, m);  // this mode's terms, stride M (model_intensity sums them)
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = F[g.e(a)];
    s256_line<true, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      F[g.e(a)] = x[a];
      I[(size_t)g.e(a) * M] = abs2f(x[a]) * S256_INV_NN;
    }
  } else if constexpr (OP == C_MODULUS) {
    // chi_hat = (sqrt(D + c)/(sqrt(I + c) + eps) - 1) mask F (noise_models.py:47-63), then
    // its inverse column transform into chi_a.  flags bit 0: the error sum of the slot's
    // pattern (solvers.py:237), this tile's part, from the mode-0 plane into red.err
    const float2* F = (const float2*)d.farfield + plane;
    const float* D = d.patterns_s + (size_t)d.group_pos[b] * S256_NN;
    const float off = d.params[P_OFFSET], eps = d.params[P_EPS];
    float2* chi = (float2*)d.chi_a + plane;
    float e2 = 0.f;
    auto body = [&](auto isum) __attribute__((always_inline)) {
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        const int e = g.e(a);
        float t;
        const float u = farfield_factor<FM_MODULUS>(isum(e), D[e], d.mask_s[e], off, eps, 0.f, 1.f, t);
        e2 += t;
        x[a] = cscale(F[e], u);
      }
    };
    if (M <= 4) {
      const ModeSum4 ms(d, b);
      body([&](int e) { return ms(e); });
    } else {
      body([&](int e) { return model_intensity(d, b, e); });
    }
    if ((flags & 1) && m == 0) {
      const double tot = s256_block_sum<true>(e2, s);
      if (threadIdx.x == 0) d.red[red_layout(d.n_slices, d.bmax, M).err + 8 * b + tile] = tot;
    }
    s256_line<true, true>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) chi[g.e(a)] = x[a];
  } else if constexpr (OP == C_GRAD_AMPLITUDE || OP == C_GRAD_POISSON) {
    // gradient of the amplitude (noise_models.py:29-45) or Poisson (:86-102) loss w.r.t.
    // the exit wave (torch's convention, times params[P_LOSS_SCALE]), then its inverse
    // column transform into chi_a; the loss of the mode-0 planes goes to red.loss (FP64)
    constexpr int FM = OP == C_GRAD_AMPLITUDE ? FM_GRAD_AMPLITUDE : FM_GRAD_POISSON;
    const float2* F = (const float2*)d.farfield + plane;
    const float* D = d.patterns_s + (size_t)d.group_pos[b] * S256_NN;
    const float off = d.params[P_OFFSET], eps = d.params[P_EPS], var = d.params[P_VAR];
    const float gscale = d.params[P_LOSS_SCALE];
    float2* chi = (float2*)d.chi_a + plane;
    float e2 = 0.f;
    auto body = [&](auto isum) __attribute__((always_inline)) {
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        const int e = g.e(a);
        float t;
        const float u = farfield_factor<FM>(isum(e), D[e], d.mask_s[e], off, eps, var, gscale, t);
        e2 += t;
        x[a] = cscale(F[e], u);
      }
    };
    if (M <= 4) {
      const ModeSum4 ms(d, b);
      body([&](int e) { return ms(e); });
    } else {
      body([&](int e) { return model_intensity(d, b, e); });
    }
    if (m == 0) add_group_loss<FM>(d, e2, s.red);
    s256_line<true, true>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) chi[g.e(a)] = x[a];
  } else if constexpr (OP == C_GRAD_PROBE) {
    // slice 0 of the adjoint: the probe-gradient spectrum fft2(g_psi0) conj(filt_b) into the
    // farfield slot (flags bit 0; summed over positions in order by the probe finish) and
    // the position gradients Re sum conj(fft2 g_psi0) fft2(P) filt (-2 pi i k) (bit 1)
    const bool want_probe = flags & 1, want_pos = flags & 2;
    const int p = d.group_pos[b];
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    const float2* Pk = (const float2*)d.probe_k + (size_t)m * S256_NN;
    float2* Fb = (float2*)d.farfield + plane;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = Fb[g.e(a)];
    s256_line<true, false>(x, s, g.l, g.t);
    float gy = 0.f, gx = 0.f;
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int e = g.e(a);
      const float ky = d.ky_s[g.y(a)], kx = d.kx_s[g.x(a)];
      const float2 f = shift_filter(sy, sx, ky, kx);
      if (want_probe) Fb[e] = cmulc(x[a], f);
      if (want_pos) {
        const float2 t = cmulc(cmul(Pk[e], f), x[a]);  // conj(v) fft2(P) filt
        gy += TWO_PI * ky * t.y;
        gx += TWO_PI * kx * t.y;
      }
    }
    if (want_pos) {
      const double ty = s256_block_sum<true>(gy, s), tx = s256_block_sum<true>(gx, s);
      if (threadIdx.x == 0) {
        const RedLayout L = red_layout(d.n_slices, d.bmax, M);
        atomic_add_f64(&d.red[L.pos_num + 2 * b], ty * (double)S256_INV_NN);
        atomic_add_f64(&d.red[L.pos_num + 2 * b + 1], tx * (double)S256_INV_NN);
      }
    }
  } else if constexpr (OP == C_POS) {
    const int dir = flags & 1;
    float2* w = (float2*)d.farfield + plane;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = w[g.e(a)];
    s256_line<true, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const float wk = TWO_PI * (dir == 0 ? d.ky_s[g.y(a)] : d.kx_s[g.x(a)]);
      x[a] = make_float2(wk * x[a].y, -wk * x[a].x);  // v * (-2 pi i k)
    }
    s256_line<true, true>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a) w[g.e(a)] = x[a];
  } else if constexpr (OP == C_BACK_PROBE) {
    // per-position probe-update spectrum fft2(dP) conj(filt_b) (solvers.py:262-270), summed
    // over positions in order by probe_spectrum_reduce_k
    const int p = d.group_pos[b];
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    float2* Fp = (float2*)d.farfield + plane;
#pragma unroll
    for (int a = 0; a < 16; ++a) x[a] = Fp[g.e(a)];
    s256_line<true, false>(x, s, g.l, g.t);
#pragma unroll
    for (int a = 0; a < 16; ++a)
      Fp[g.e(a)] = cmulc(x[a], shift_filter(sy, sx, d.ky_s[g.y(a)], d.kx_s[g.x(a)]));
  }
}

// errors[b] = sqrt(nansum((I - D)^2)) over the slot's pattern (solvers.py:237; unmasked):
// the 8 column tiles' FP64 partials that C_MODULUS left in red.err, added in tile order.
__global__ void __launch_bounds__(64) s256_errors_k(phx_group_desc d) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= d.bmax || !slot_active(d, b)) return;
  const double* part = d.red + red_layout(d.n_slices, d.bmax, d.n_modes).err + 8 * b;
  double e2 = 0.0;
  for (int t = 0; t < S256_COL_TILES; ++t) e2 += part[t];
  d.errors[b] = (float)sqrt(e2);
}

// ----- host launchers (n = 256) -----
inline dim3 s256_row_grid(const phx_group_desc& d, bool per_mode) {
  return dim3((unsigned)(S256_ROW_TILES * d.n_modes), per_mode ? 1u : (unsigned)d.bmax);
}
inline dim3 s256_col_grid(const phx_group_desc& d, bool per_mode) {
  return dim3((unsigned)(S256_COL_TILES * d.n_modes), per_mode ? 1u : (unsigned)d.bmax);
}
#define S256_ROWS(OP, PM, I, F) \
  hipLaunchKernelGGL((s256_rows_k<OP>), s256_row_grid(d, PM), dim3(256), 0, stream, d, (I), (F))
#define S256_COLS(OP, PM, I, F) \
  hipLaunchKernelGGL((s256_cols_k<OP>), s256_col_grid(d, PM), dim3(512), 0, stream, d, (I), (F))
#define S256_ERRORS() \
  hipLaunchKernelGGL(s256_errors_k, dim3((unsigned)((d.bmax + 63) / 64)), dim3(64), 0, stream, d)

inline int s256_probe_spectra(const phx_group_desc& d, hipStream_t stream) {
  S256_ROWS(R_PROBE_FWD, true, 0, 0);
  S256_COLS(C_PROBE_FWD, true, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

inline int s256_shift_probes(const phx_group_desc& d, int propagate, hipStream_t stream) {
  S256_COLS(C_SHIFT, false, 0, 0);
  S256_ROWS(R_SHIFT_FINISH, false, 0, propagate ? 1 : 0);
  if (propagate) {
    S256_COLS(C_PROPAGATE, false, 0, 0);
    S256_ROWS(R_WAVE_FINISH, false, 0, 0);
  }
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

inline int s256_forward_slice(const phx_group_desc& d, int i, hipStream_t stream) {
  S256_ROWS(R_SLICE_START, false, i, 0);
  S256_COLS(C_PROPAGATE, false, i, 0);
  S256_ROWS(R_WAVE_FINISH, false, i, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// farfield = FFT(psi[S-1] O_{S-1}) and each mode's |farfield|^2 / n^2 term of the intensity
inline int s256_exit_spectra(const phx_group_desc& d, hipStream_t stream) {
  S256_ROWS(R_EXIT, false, 0, 0);
  S256_COLS(C_EXIT, false, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

inline int s256_farfield(const phx_group_desc& d, int calc_error, hipStream_t stream) {
  S256_ROWS(R_EXIT, false, 0, 0);
  S256_COLS(C_EXIT, false, 0, 0);
  S256_COLS(C_MODULUS, false, 0, calc_error ? 1 : 0);
  if (calc_error) S256_ERRORS();
  S256_ROWS(R_CHI_FINISH, false, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// The forward model of a group with each slice step's finishing row pass folded into the
// next step's first one: psi_0, then per slice a column pass (propagator) and one row pass
// (finish psi_{i+1}, store it, transmit, forward rows), then the exit spectra: 2S + 1
// passes instead of 3S + 1.
inline int s256_group_forward(const phx_group_desc& d, hipStream_t stream) {
  const int S = d.n_slices;
  S256_COLS(C_SHIFT, false, 0, 0);
  if (S == 1) {
    S256_ROWS(R_SHIFT_FINISH, false, 0, 0);
    S256_ROWS(R_EXIT, false, 0, 0);
  } else {
    S256_ROWS(R_SHIFT_FINISH, false, 0, 1);
    for (int i = 0; i < S - 1; ++i) {
      S256_COLS(C_PROPAGATE, false, i, 0);
      S256_ROWS(R_SLICE_START, false, i + 1, 8);  // finish psi_{i+1}, then its step
    }
  }
  S256_COLS(C_EXIT, false, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// The far-field modulus update alone: chi_a = ifft2(u F), errors when calc_error.
inline int s256_farfield_update(const phx_group_desc& d, int calc_error, hipStream_t stream) {
  S256_COLS(C_MODULUS, false, 0, calc_error ? 1 : 0);
  if (calc_error) S256_ERRORS();
  S256_ROWS(R_CHI_FINISH, false, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// The LSQ-ML backward sweep with each step's finishing row pass folded into the next
// step's first one: the modulus column pass, then per slice one row pass (finish chi_i,
// store it, the step-size sums, dP, forward rows) and a column pass (conj propagator),
// and at slice 0 the probe spectra: 2S + 1 passes (+1 with a probe update) instead of 3S + 2.
inline int s256_lsqml_backward_sweep(const phx_group_desc& d, int upd_obj, int upd_probe, int calc_error,
                                     hipStream_t stream) {
  const int S = d.n_slices;
  S256_COLS(C_MODULUS, false, 0, calc_error ? 1 : 0);
  if (calc_error) S256_ERRORS();
  for (int i = S - 1; i >= 0; --i) {
    const bool probe = i == 0 && upd_probe;
    S256_ROWS(R_BACK, false, i, (upd_obj ? 1 : 0) | (probe ? 2 : 0) | 8);
    if (i > 0) S256_COLS(C_PROPAGATE, false, i - 1, 1);
    else if (probe) S256_COLS(C_BACK_PROBE, false, 0, 0);
  }
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

inline int s256_lsqml_backward(const phx_group_desc& d, int i, int upd_obj, int upd_probe, hipStream_t stream) {
  const bool probe = i == 0 && upd_probe;
  S256_ROWS(R_BACK, false, i, (upd_obj ? 1 : 0) | (probe ? 2 : 0));
  if (i > 0) {
    // chi_{i-1} = ifft2(fft2(dP) conj(p_{i-1})), in chi_out_buf(i)
    S256_COLS(C_PROPAGATE, false, i - 1, 1);
    S256_ROWS(R_BACK_FINISH, false, i, 0);
  } else if (probe) {
    S256_COLS(C_BACK_PROBE, false, 0, 0);
  }
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// ePIE backward step at slice i (after epie_maxima_k): back-propagation (i > 0) or the
// per-position probe spectrum (i == 0 with a probe update); at slice 0 without one there
// is nothing to do (the object update is gathered by epie_object_update_k).
inline int s256_epie_backward(const phx_group_desc& d, int i, int upd_probe, hipStream_t stream) {
  if (i == 0 && !upd_probe) return PHX_OK;
  S256_ROWS(R_EPIE_BACK, false, i, 0);
  if (i > 0) {
    S256_COLS(C_PROPAGATE, false, i - 1, 1);
    S256_ROWS(R_BACK_FINISH, false, i, 0);
  } else {
    S256_COLS(C_BACK_PROBE, false, 0, 0);
  }
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// ePIE probe step after the ordered sum of the per-position spectra into d_probe
inline int s256_epie_probe_apply(const phx_group_desc& d, hipStream_t stream) {
  S256_COLS(C_PROBE_INV, true, 0, 0);
  S256_ROWS(R_EPIE_PROBE, true, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// Gradient engine: the loss gradient w.r.t. the exit wave into chi_a and the group loss,
// from the exit spectra a forward launch left
inline int s256_farfield_gradient_update(const phx_group_desc& d, int poisson, hipStream_t stream) {
  if (poisson) S256_COLS(C_GRAD_POISSON, false, 0, 0);
  else S256_COLS(C_GRAD_AMPLITUDE, false, 0, 0);
  S256_ROWS(R_CHI_FINISH, false, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// The same after the exit wave's own passes
inline int s256_farfield_gradient(const phx_group_desc& d, int poisson, hipStream_t stream) {
  S256_ROWS(R_EXIT, false, 0, 0);
  S256_COLS(C_EXIT, false, 0, 0);
  if (poisson) S256_COLS(C_GRAD_POISSON, false, 0, 0);
  else S256_COLS(C_GRAD_AMPLITUDE, false, 0, 0);
  S256_ROWS(R_CHI_FINISH, false, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// Gradient engine adjoint step at slice i (the object gradient is gathered afterwards from
// the wave gradients the steps leave in chi_in_buf): back-propagation (i > 0), or at
// slice 0 the probe-gradient spectra and position gradients.
inline int s256_gradient_backward(const phx_group_desc& d, int i, int want_probe, int want_pos,
                                  hipStream_t stream) {
  if (i > 0) {
    S256_ROWS(R_BACK, false, i, 0);
    S256_COLS(C_PROPAGATE, false, i - 1, 1);
    S256_ROWS(R_BACK_FINISH, false, i, 0);
  } else if (want_probe || want_pos) {
    S256_ROWS(R_BACK, false, 0, 4);
    S256_COLS(C_GRAD_PROBE, false, 0, (want_probe ? 1 : 0) | (want_pos ? 2 : 0));
  }
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// grad_probe = ifft2(d_probe) after the ordered sum of the per-position spectra
inline int s256_gradient_probe_finish(const phx_group_desc& d, hipStream_t stream) {
  S256_COLS(C_PROBE_INV, true, 0, 0);
  S256_ROWS(R_GRAD_PROBE, true, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// LSQ-ML position step in direction dir.  Its work plane is the farfield slot, so at this
// size it runs before the slice-0 backward step writes the probe spectra there (it reads
// only chi_0, psi_0 and the pre-update slice-0 cutout).
inline int s256_position_gradient(const phx_group_desc& d, int dir, hipStream_t stream) {
  S256_ROWS(R_POS_START, false, 0, 0);
  S256_COLS(C_POS, false, 0, dir);
  S256_ROWS(R_POS_FINISH, false, 0, dir);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

inline int s256_lsqml_probe_apply(const phx_group_desc& d, hipStream_t stream) {
  S256_COLS(C_PROBE_INV, true, 0, 0);
  S256_ROWS(R_PROBE_UPDATE, true, 0, 0);
  S256_COLS(C_PROBE_SPEC, true, 0, 0);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

#undef S256_ROWS
#undef S256_COLS
#undef S256_ERRORS
