// Batched 2-D complex FFT of n x n planes (n = 16, 32, 64, 128, 192) for gfx950, one
// plane per workgroup of L n threads.
//
// Layout.  A group of L lanes (L = 8 for n = 128, 4 otherwise) owns one row (row
// phase) or one column (column phase) of the plane; lane q holds the E = n/L points
// with index L j + q, j < E.  A length-n transform along the group is a four-step
// n = E * L decomposition:
//   1. in-register DFT_E over j (radix-2, compile-time twiddles; E = 48 as 3 x 16),
//   2. twiddle W_n^(q k1) (table in LDS),
//   3. DFT_L across the lanes: for L = 4 the DFT_4 of each quad as two radix-2
//      exchanges over DPP permutes; for L = 8 an exchange through LDS within the wave,
//      then two in-register DFT_8s per lane.
// The transform leaves each lane's results at "scrambled" positions p = L k + q
// (register k of lane q); scr_to_freq gives the frequency held at p.  The inverse runs
// the same steps backwards and returns natural order.  At n = 128 the 8-lane layout
// gives 1024 threads per plane: with one plane per CU (its LDS), that is 4 waves per
// SIMD to hide each phase's latency instead of 2.
//
// A 2-D transform is: rows in registers -> transpose through LDS -> columns in
// registers.  At n = 192 the complex plane (288 KB) exceeds the 160 KB LDS, so the
// transpose moves the real parts and then the imaginary parts through one 192 x 192
// float plane (144 KB).
// The LDS plane is padded (and for L = 8 XOR-swizzled, Plane::row_base / col_base) so that both the
// row-phase accesses (s[r][L k + q]) and the column-phase ones (s[L j + q][c]) are
// bank-conflict free.  A forward transform, a pointwise multiply in the (scrambled)
// frequency domain and an inverse transform therefore cost two LDS round trips in total
// (six at n = 128, with the in-wave exchanges).
//
// Transforms are unnormalised: forward uses exp(-2 pi i nk/n), inverse exp(+2 pi i nk/n),
// so inverse(forward(a)) = n^2 a; callers scale by 1/n^2.  The (i)fftshift of the centred
// real-space convention cancels in every fft2 -> pointwise -> ifft2 product the engines
// form, so it is not applied here.
#pragma once
#include "phx_common.h"

// Points per lane whose loads are issued together in a load phase (a scheduling
// barrier after every PHX_LOAD_PACE points).
#ifndef PHX_LOAD_PACE
#define PHX_LOAD_PACE 8
#endif

namespace phx {

// exp(-2 pi i k / 32), k in [0, 32)
__device__ __forceinline__ float2 w32(int k) {
  constexpr float C1 = 0.98078528040323044f, C2 = 0.92387953251128676f, C3 = 0.83146961230254524f,
                  C4 = 0.70710678118654752f, C5 = 0.55557023301960222f, C6 = 0.38268343236508977f,
                  C7 = 0.19509032201612826f;
  constexpr float c[32] = {1.f, C1, C2, C3, C4, C5, C6, C7, 0.f, -C7, -C6, -C5, -C4, -C3, -C2, -C1,
                           -1.f, -C1, -C2, -C3, -C4, -C5, -C6, -C7, 0.f, C7, C6, C5, C4, C3, C2, C1};
  // -sin(2 pi k/32) = c[(k + 8) % 32] * -1 ... written out for clarity
  constexpr float s[32] = {0.f, -C7, -C6, -C5, -C4, -C3, -C2, -C1, -1.f, -C1, -C2, -C3, -C4, -C5, -C6, -C7,
                           0.f, C7, C6, C5, C4, C3, C2, C1, 1.f, C1, C2, C3, C4, C5, C6, C7};
  return make_float2(c[k], s[k]);
}

// Radix-2 butterfly a = u + W v, b = u - W v with W = W_32^k (k compile-time after
// unrolling).  Trivial factors need no multiply.  The others use the tangent form of the
// rotation: W v = g p with p = v + i (h/g) v-type terms (g the larger of |Re W|, |Im W|),
// so the product and the add/subtract take 2 + 4 FMAs instead of a complex multiply
// (4) plus an add and a subtract (4).
__device__ __forceinline__ void bfly32(float2& a, float2& b, float2 u, float2 v, int k) {
  if (k == 0) {
    a = cadd(u, v);
    b = csub(u, v);
    return;
  }
  if (k == 8 || k == 24) {  // W = -i or +i
    const float2 t = k == 8 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
    a = cadd(u, t);
    b = csub(u, t);
    return;
  }
  if (k == 16) {
    a = csub(u, v);
    b = cadd(u, v);
    return;
  }
  const float2 w = w32(k);
  float g;
  float2 p;
  if (fabsf(w.x) >= fabsf(w.y)) {  // W v = w.x (v.x - r v.y, v.y + r v.x), r = w.y / w.x
    const float r = w.y / w.x;
    g = w.x;
    p = make_float2(fmaf(-r, v.y, v.x), fmaf(r, v.x, v.y));
  } else {  // W v = w.y (r v.x - v.y, v.x + r v.y), r = w.x / w.y
    const float r = w.x / w.y;
    g = w.y;
    p = make_float2(fmaf(r, v.x, -v.y), fmaf(r, v.y, v.x));
  }
  a = make_float2(fmaf(g, p.x, u.x), fmaf(g, p.y, u.y));
  b = make_float2(fmaf(-g, p.x, u.x), fmaf(-g, p.y, u.y));
}

template <int R> struct Log2;
template <> struct Log2<1> { static constexpr int v = 0; };
template <> struct Log2<2> { static constexpr int v = 1; };
template <> struct Log2<4> { static constexpr int v = 2; };
template <> struct Log2<8> { static constexpr int v = 3; };
template <> struct Log2<16> { static constexpr int v = 4; };
template <> struct Log2<32> { static constexpr int v = 5; };

// exp(-2 pi i e / 48), e in [0, 31) (the twiddles of the 3 x 16 split below)
__device__ __forceinline__ float2 w48(int e) {
  constexpr float c[31] = {
      1.f, 0.99144486137381038f, 0.96592582628906831f, 0.92387953251128674f, 0.86602540378443865f,
      0.79335334029123517f, 0.70710678118654752f, 0.60876142900872066f, 0.5f, 0.38268343236508977f,
      0.25881904510252074f, 0.13052619222005159f, 0.f, -0.13052619222005159f, -0.25881904510252074f,
      -0.38268343236508977f, -0.5f, -0.60876142900872066f, -0.70710678118654752f, -0.79335334029123517f,
      -0.86602540378443865f, -0.92387953251128674f, -0.96592582628906831f, -0.99144486137381038f, -1.f,
      -0.99144486137381038f, -0.96592582628906831f, -0.92387953251128674f, -0.86602540378443865f,
      -0.79335334029123517f, -0.70710678118654752f};
  constexpr float s[31] = {
      0.f, -0.13052619222005159f, -0.25881904510252074f, -0.38268343236508977f, -0.5f,
      -0.60876142900872066f, -0.70710678118654752f, -0.79335334029123517f, -0.86602540378443865f,
      -0.92387953251128674f, -0.96592582628906831f, -0.99144486137381038f, -1.f, -0.99144486137381038f,
      -0.96592582628906831f, -0.92387953251128674f, -0.86602540378443865f, -0.79335334029123517f,
      -0.70710678118654752f, -0.60876142900872066f, -0.5f, -0.38268343236508977f, -0.25881904510252074f,
      -0.13052619222005159f, 0.f, 0.13052619222005159f, 0.25881904510252074f, 0.38268343236508977f, 0.5f,
      0.60876142900872066f, 0.70710678118654752f};
  return make_float2(c[e], s[e]);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&x)[R]);

// In-register length-48 DFT (natural order in and out): 48 = 3 x 16, decimation in
// time: X[k + 16 k2] = sum_b W_3^(b k2) W_48^(b k) Y_b[k], Y_b = DFT_16 of x[3 a + b].
template <bool INV>
__device__ __forceinline__ void dft48(float2 (&x)[48]) {
  float2 y[3][16];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 16; ++a) y[b][a] = x[3 * a + b];
#pragma unroll
  for (int b = 0; b < 3; ++b) dft<16, INV>(y[b]);
#pragma unroll
  for (int b = 1; b < 3; ++b)
#pragma unroll
    for (int k = 1; k < 16; ++k) y[b][k] = INV ? cmulc(y[b][k], w48(b * k)) : cmul(y[b][k], w48(b * k));
  constexpr float h = 0.86602540378443865f;  // sqrt(3) / 2
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    // DFT-3: X0 = a + s, X1,2 = (a - s/2) -+ i h d (forward; +- in the inverse)
    const float2 a = y[0][k], sm = cadd(y[1][k], y[2][k]), d = csub(y[1][k], y[2][k]);
    const float2 t = make_float2(fmaf(-0.5f, sm.x, a.x), fmaf(-0.5f, sm.y, a.y));
    const float2 r = INV ? make_float2(-h * d.y, h * d.x) : make_float2(h * d.y, -h * d.x);
    x[k] = cadd(a, sm);
    x[k + 16] = cadd(t, r);
    x[k + 32] = csub(t, r);
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft_pow2(float2 (&x)[R]);

// In-register length-R DFT (R = 2..32 or 48), natural order in and out.
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
  if constexpr (R == 48) {
    dft48<INV>(x);
  } else {
    dft_pow2<R, INV>(x);
  }
}

// radix-2, R = 2..32
template <int R, bool INV>
__device__ __forceinline__ void dft_pow2(float2 (&x)[R]) {
  constexpr int LG = Log2<R>::v;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int j = 0;
#pragma unroll
    for (int b = 0; b < LG; ++b) j |= ((i >> b) & 1) << (LG - 1 - b);
    if (j > i) {
      const float2 t = x[i];
      x[i] = x[j];
      x[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        int k = j * (32 / len);
        if (INV) k = (32 - k) & 31;
        bfly32(x[i + j], x[i + j + half], x[i + j], x[i + j + half], k);
      }
    }
  }
}

// Quad permutes (DPP quad_perm): value of lane q ^ 1 and of lane q ^ 2.
__device__ __forceinline__ float dpp_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ float2 quad_xor1(float2 v) { return make_float2(dpp_xor1(v.x), dpp_xor1(v.y)); }
__device__ __forceinline__ float2 quad_xor2(float2 v) { return make_float2(dpp_xor2(v.x), dpp_xor2(v.y)); }

// Lanes per row (row phase) or column (column phase): 8 at n = 128 (1024 threads: 4
// waves per SIMD for the one plane a CU holds), 4 otherwise (n = 192: 768 threads of 48
// points each).
template <int N>
struct PlaneLanes {
  static constexpr int value = N == 128 ? 8 : 4;
};

// Frequency index held at scrambled position p = L k + q of an N-point axis (register k
// of lane q after a forward transform).  4 lanes: X[k + (N/4) bitrev2(q)].  8 lanes
// (N = 128): register k = 2 r + s of lane q holds X[2 q + s + 16 r].
template <int N>
__host__ __device__ __forceinline__ int scr_to_freq(int p) {
  if (N == 256) return p;  // 256 x 256 planes (split over launches) keep natural order
  constexpr int L = PlaneLanes<N>::value;
  const int q = p % L, k = p / L;
  if (L == 8) return 2 * q + (k & 1) + 16 * (k >> 1);
  return k + (N / 4) * (((q & 1) << 1) | (q >> 1));
}

template <int N>
struct Plane {
  static constexpr int QL = PlaneLanes<N>::value;  // lanes per row / column
  static constexpr int E = N / QL;      // points per lane
  static constexpr int T = QL * N;      // threads per workgroup
  // LDS row stride (float2).  4 lanes: n + 4 makes both transpose directions bank-
  // conflict free.  8 lanes: no plain stride does (one direction stays 2-way), so n + 8
  // plus an XOR of column bit 2 with row bit 2 (checked exhaustively for every half-wave
  // of both access patterns: 64 distinct banks each).
  // SPLIT (n = 192): the plane is transposed in two passes, real parts then imaginary
  // parts, through an n x n float plane without padding.  Float element (r, c) lives at
  // r n + (c ^ swz(r)), swz(r) = 4 b2 + 8 (b1 ^ b3) + 16 b0 + 32 b1 (b_i = bit i of r):
  // for every wave and register, the 64 lanes of a row-phase access (16 rows x 4 lanes)
  // and of a column-phase access (16 columns x 4 rows) hit 64 distinct banks mod 64,
  // and each half-wave 32 distinct banks mod 32 (checked exhaustively).  swz only
  // touches column bits 2-5, so it stays inside the aligned 64-column block.
  static constexpr bool SPLIT = N == 192;
  static constexpr int LD = SPLIT ? N : (QL == 8 ? N + 8 : N + 4);
  static constexpr int LDS_FLOAT2 = SPLIT ? N * N / 2 : N * LD;
  // element (row, col) lives at row * LD + (col ^ sw(row)); for the two transpose
  // patterns the swizzle is a per-lane constant, so each is one base offset per lane
  // plus compile-time steps
  __device__ __forceinline__ int row_base() const {  // (row rc, col QL k + q): + QL k
    return rc * LD + (QL == 8 ? (q ^ (((rc >> 2) & 1) << 2)) : q);
  }
  __device__ __forceinline__ int col_base() const {  // (row QL j + q, col rc): + QL LD j
    return q * LD + (QL == 8 ? (rc ^ (((q >> 2) & 1) << 2)) : rc);
  }
  float2* s;         // N x LD in LDS (SPLIT: N x N floats)
  const float2* tw;  // twiddles in LDS: tw[QL k + q] = W_N^(q k)
  int rc;            // row (row phase) or column (column phase) owned by this lane group
  int q;             // lane within the group

  __device__ __forceinline__ Plane(float2* s_, const float2* tw_) : s(s_), tw(tw_) {
    rc = threadIdx.x / QL;
    q = threadIdx.x % QL;
  }

  // Makes rc and q opaque for one loop iteration: called at the top of a loop over
  // slices, it keeps the compiler from hoisting every point's address arithmetic out of
  // the loop (dozens of registers live across all transforms).
  __device__ __forceinline__ void refresh() { __asm__ volatile("" : "+v"(rc), "+v"(q)); }

  // ----- one 1-D transform along the lane group -----
  // After the in-lane DFT_E and the twiddle W_N^(q k), the DFT across the L lanes:
  //  * 4 lanes: two radix-2 exchanges over DPP quad_perm (partner q^2, then q^1), each
  //    one FMA per component with a per-lane sign (+-1 is exact), and the one rotation
  //    of the radix-2x2 split (by -i on lane 3 in the forward, +i in the inverse,
  //    applied before the exchange that reads it) as a select;
  //  * 8 lanes (N = 128, E = 16): an exchange through LDS inside the wave, after which
  //    lane q holds, for k1 = 2q and 2q+1, the values of all 8 lanes, and two in-lane
  //    DFT-8s finish the transform (fewer VALU instructions than three DPP stages).
  // No lane-dependent branches: a group's lanes run the same instructions.
  //
  // Exchange slots (8 lanes).  A wave uses the LDS rows of its own 8 row groups,
  // 8 (rc >> 3) .. + 7 (1088 float2 from a multiple of 32): value (k1 = 2h + s, of lane
  // qq) of lane group g = rc & 7 sits at 8 (g & 3) + 526 (g >> 2) + qq + 33 h + 263 s.
  // One half-wave is 4 groups (same g >> 2); writing all qq for one k1, or reading all
  // h for one (qq, s), its 32 lanes land on 8 (g & 3) + (qq or h) + const mod 32: 32
  // distinct float2 banks.  Both forms are a per-lane base plus compile-time offsets.
  // In the row phase no other wave touches these rows; in the column phase other waves
  // read them during the transpose, so the forward column pass waits at a barrier first
  // (an inverse column pass always follows a forward one or starts a kernel).
  __device__ __forceinline__ int xbase() const {
    const int g = rc & 7;
    return (rc >> 3) * 8 * LD + 8 * (g & 3) + 526 * (g >> 2);
  }
  template <bool COL>
  __device__ __forceinline__ void fwd(float2 (&x)[E]) {
    dft<E, false>(x);
    // tw[q] = W^0 = 1 exactly, so k = 0 needs no multiply
    const float2* twq = tw + q;
#pragma unroll
    for (int k = 1; k < E; ++k) x[k] = cmul(x[k], twq[QL * k]);
    if constexpr (QL == 8) {
      if (COL) __syncthreads();
      const int wb = xbase() + q, rb = xbase() + 33 * q;
#pragma unroll
      for (int k = 0; k < E; ++k) s[wb + 33 * (k >> 1) + 263 * (k & 1)] = x[k];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float2 y[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) y[t] = s[rb + t + 263 * h];
        dft<8, false>(y);
#pragma unroll
        for (int r = 0; r < 8; ++r) x[2 * r + h] = y[r];
      }
    } else {
      const int qq = q & 3;
      const float s1 = qq < 2 ? 1.f : -1.f;   // y = x + p (qq < 2), p - x (qq >= 2)
      const float s2 = (qq & 1) ? -1.f : 1.f; // out = y + p (even), p - y (odd)
      const bool r3 = qq == 3;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const float2 p = quad_xor2(x[k]);
        const float yr = fmaf(s1, x[k].x, p.x), yi = fmaf(s1, x[k].y, p.y);
        const float2 y = make_float2(r3 ? yi : yr, r3 ? -yr : yi);  // * W4^1 = -i
        const float2 p2 = quad_xor1(y);
        x[k] = make_float2(fmaf(s2, y.x, p2.x), fmaf(s2, y.y, p2.y));
      }
    }
  }
  template <bool COL>
  __device__ __forceinline__ void inv(float2 (&x)[E]) {
    if constexpr (QL == 8) {
      const int wb = xbase() + 33 * q, rb = xbase() + q;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float2 y[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) y[r] = x[2 * r + h];
        dft<8, true>(y);
#pragma unroll
        for (int t = 0; t < 8; ++t) s[wb + t + 263 * h] = y[t];
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < E; ++k) x[k] = s[rb + 33 * (k >> 1) + 263 * (k & 1)];
    } else {
      const int qq = q & 3;
      const float s1 = qq < 2 ? 1.f : -1.f;
      const float s2 = (qq & 1) ? -1.f : 1.f;
      const bool r3 = qq == 3;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const float2 p = quad_xor1(x[k]);
        const float ar = fmaf(s2, x[k].x, p.x), ai = fmaf(s2, x[k].y, p.y);
        const float2 a = make_float2(r3 ? -ai : ar, r3 ? ar : ai);  // * conj(W4^1) = +i
        const float2 p2 = quad_xor2(a);
        x[k] = make_float2(fmaf(s1, a.x, p2.x), fmaf(s1, a.y, p2.y));
      }
    }
    const float2* twq = tw + q;
#pragma unroll
    for (int k = 1; k < E; ++k) x[k] = cmulc(x[k], twq[QL * k]);
    dft<E, true>(x);
  }

  // ----- transposes through LDS -----
  // SPLIT offsets (floats).  Row phase, register k = 16 kh + kl: element (rc, 4 k + q) at
  // rc n + q + 64 kh + (4 kl ^ swz(rc)); rc n + q has bits 2-5 clear, so with
  // srow_base = (rc n + q) | swz(rc) it is (srow_base ^ 4 kl) + 64 kh: 16 XORs serve
  // all 48 registers.  Column phase, register j: element (4 j + q, rc); swz(4 j + q) =
  // (16 b0(q) + 40 b1(q)) ^ 4 (j & 3), so it is (scol_base ^ 4 (j & 3)) + 4 n j with
  // scol_base = q n + (rc ^ (16 b0(q) + 40 b1(q))).
  __device__ __forceinline__ static int swz(int r) {
    return 4 * ((r >> 2) & 1) + 8 * (((r >> 1) ^ (r >> 3)) & 1) + 16 * (r & 1) + 32 * ((r >> 1) & 1);
  }
  __device__ __forceinline__ int srow_base() const { return (rc * N + q) | swz(rc); }
  __device__ __forceinline__ int scol_base() const { return q * N + (rc ^ (16 * (q & 1) + 40 * (q >> 1))); }
  __device__ __forceinline__ static int srow(int base, int k) { return (base ^ (4 * (k & 15))) + 64 * (k >> 4); }
  __device__ __forceinline__ static int scol(int base, int j) { return (base ^ (4 * (j & 3))) + 4 * N * j; }
  // Each pass starts from opaque copies of the bases, so the compiler computes the XORed
  // addresses inside the pass instead of keeping 20 of them live across the transforms.
  __device__ __forceinline__ static int opaque(int v) {
    __asm__ volatile("" : "+v"(v));
    return v;
  }

  // row phase (x[k] at scrambled column QL k + q of row rc) -> column phase (x[j] at row
  // QL j + q of column rc)
  __device__ __forceinline__ void rows_to_cols(float2 (&x)[E]) {
    if constexpr (SPLIT) {
      // the column reads of a previous transpose may still be in flight: barrier first;
      // each pass overwrites one component of x in place after all of it was written
      float* sf = reinterpret_cast<float*>(s);
      const int rb = srow_base(), cb = scol_base();
      __syncthreads();
      {
        const int b = opaque(rb);
#pragma unroll
        for (int k = 0; k < E; ++k) sf[srow(b, k)] = x[k].x;
      }
      __syncthreads();
      {
        const int b = opaque(cb);
#pragma unroll
        for (int j = 0; j < E; ++j) x[j].x = sf[scol(b, j)];
      }
      __syncthreads();
      {
        const int b = opaque(rb);
#pragma unroll
        for (int k = 0; k < E; ++k) sf[srow(b, k)] = x[k].y;
      }
      __syncthreads();
      {
        const int b = opaque(cb);
#pragma unroll
        for (int j = 0; j < E; ++j) x[j].y = sf[scol(b, j)];
      }
      return;
    }
    // The write goes to this wave's own rows.  With 8 lanes, every access of them by
    // another wave (a transpose's column reads or writes) is followed by a barrier before
    // the owner touches them again (the forward column pass's, or the one inside
    // cols_to_rows), so none is needed here; with 4 lanes the column reads of a previous
    // forward transform may still be in flight.
    if (QL != 8) __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) s[row_base() + QL * k] = x[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = s[col_base() + QL * LD * j];
  }
  // column phase (x[j] at row QL j + q of column rc) -> row phase (x[k] at column QL k + q
  // of row rc)
  __device__ __forceinline__ void cols_to_rows(float2 (&x)[E]) {
    if constexpr (SPLIT) {
      float* sf = reinterpret_cast<float*>(s);
      const int rb = srow_base(), cb = scol_base();
      __syncthreads();
      {
        const int b = opaque(cb);
#pragma unroll
        for (int j = 0; j < E; ++j) sf[scol(b, j)] = x[j].x;
      }
      __syncthreads();
      {
        const int b = opaque(rb);
#pragma unroll
        for (int k = 0; k < E; ++k) x[k].x = sf[srow(b, k)];
      }
      __syncthreads();
      {
        const int b = opaque(cb);
#pragma unroll
        for (int j = 0; j < E; ++j) sf[scol(b, j)] = x[j].y;
      }
      __syncthreads();
      {
        const int b = opaque(rb);
#pragma unroll
        for (int k = 0; k < E; ++k) x[k].y = sf[srow(b, k)];
      }
      return;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) s[col_base() + QL * LD * j] = x[j];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) x[k] = s[row_base() + QL * k];
  }

  // ----- composites -----
  // natural-order source src(y, x) -> registers (row phase).  Points are taken in groups
  // of PHX_LOAD_PACE: the loads of a group are issued together, then pin() runs (a
  // kernel passes one that fixes its running sums in place with an empty asm, so the
  // compiler can neither sink the sums past the transform nor wait point by point).
  template <class Src, class Pin>
  __device__ __forceinline__ void load_rows(Src src, float2 (&x)[E], Pin pin) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      x[j] = src(rc, QL * j + q);
      if ((j % PHX_LOAD_PACE) == PHX_LOAD_PACE - 1 || j == E - 1) {
        pin();
        // the group's global loads first, then its arithmetic
        __builtin_amdgcn_sched_group_barrier(0x0020, 4 * PHX_LOAD_PACE, 0);
        __builtin_amdgcn_sched_group_barrier(0x0002, 1024, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  template <class Src>
  __device__ __forceinline__ void load_rows(Src src, float2 (&x)[E]) {
    load_rows(src, x, [] {});
  }
  // x[j] = f(y, x, x[j]) over the row-phase points (natural order), in load groups
  template <class F, class Pin>
  __device__ __forceinline__ void rows_apply(float2 (&x)[E], F f, Pin pin) {
#pragma unroll
    for (int j = 0; j < E; ++j) {
      x[j] = f(rc, QL * j + q, x[j]);
      if ((j % PHX_LOAD_PACE) == PHX_LOAD_PACE - 1 || j == E - 1) {
        pin();
        __builtin_amdgcn_sched_group_barrier(0x0020, 4 * PHX_LOAD_PACE, 0);
        __builtin_amdgcn_sched_group_barrier(0x0002, 1024, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // x[k] = f(py, px, x[k]) over the column-phase points (scrambled spectrum), in load groups
  template <class F, class Pin>
  __device__ __forceinline__ void cols_apply(float2 (&x)[E], F f, Pin pin) {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      x[k] = f(QL * k + q, rc, x[k]);
      if ((k % PHX_LOAD_PACE) == PHX_LOAD_PACE - 1 || k == E - 1) {
        pin();
        __builtin_amdgcn_sched_group_barrier(0x0020, 4 * PHX_LOAD_PACE, 0);
        __builtin_amdgcn_sched_group_barrier(0x0002, 1024, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // registers (row phase, natural) -> sink dst(y, x, v)
  template <class Dst>
  __device__ __forceinline__ void store_rows(Dst dst, float2 (&x)[E]) {
#pragma unroll
    for (int j = 0; j < E; ++j) dst(rc, QL * j + q, x[j]);
  }
  // row-phase natural data -> column-phase scrambled spectrum (py = QL k + q, px = rc)
  __device__ __forceinline__ void forward_regs(float2 (&x)[E]) {
    fwd<false>(x);
    rows_to_cols(x);
    fwd<true>(x);
  }
  // column-phase scrambled spectrum -> row-phase natural data
  __device__ __forceinline__ void inverse_regs(float2 (&x)[E]) {
    inv<true>(x);
    cols_to_rows(x);
    inv<false>(x);
  }
  // forward transform of src; fn(py, px, v&) sees every scrambled spectrum value
  template <class Src, class Fn, class Pin>
  __device__ __forceinline__ void forward(Src src, Fn fn, Pin pin) {
    float2 x[E];
    load_rows(src, x, pin);
    forward_regs(x);
#pragma unroll
    for (int k = 0; k < E; ++k) fn(QL * k + q, rc, x[k]);
  }
  template <class Src, class Fn>
  __device__ __forceinline__ void forward(Src src, Fn fn) {
    forward(src, fn, [] {});
  }
  // inverse transform of the scrambled spectrum src(py, px) into dst(y, x, v)
  template <class Src, class Dst, class Pin>
  __device__ __forceinline__ void inverse(Src src, Dst dst, Pin pin) {
    float2 x[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      x[k] = src(QL * k + q, rc);
      if ((k % PHX_LOAD_PACE) == PHX_LOAD_PACE - 1 || k == E - 1) {
        pin();
        __builtin_amdgcn_sched_group_barrier(0x0020, 4 * PHX_LOAD_PACE, 0);
        __builtin_amdgcn_sched_group_barrier(0x0002, 1024, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    inverse_regs(x);
    store_rows(dst, x);
  }
  template <class Src, class Dst>
  __device__ __forceinline__ void inverse(Src src, Dst dst) {
    inverse(src, dst, [] {});
  }
  // in-register forward -> pointwise mul(py, px, v) -> inverse
  // Column-phase waves whose columns are all zero in a convolution's multiplier (bit w of
  // zero_waves for wave w: the host derives it from the band-limited propagators) skip
  // the forward column pass, the product and the inverse column pass: their product is
  // zero, and so is its inverse column transform (exactly, for finite data).  The test
  // is wave-uniform (readfirstlane), and a skipping wave takes the one barrier the
  // forward column pass would.
  __device__ __forceinline__ bool wave_is_zero(unsigned zero_waves) const {
    return (zero_waves >> __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) & 1u;
  }
  __device__ __forceinline__ void zero_column_pass(float2 (&x)[E]) {
    if (QL == 8) __syncthreads();
#pragma unroll
    for (int k = 0; k < E; ++k) x[k] = make_float2(0.f, 0.f);
  }
  // in-register forward -> pointwise mul(py, px, v) -> inverse
  template <class Mul>
  __device__ __forceinline__ void convolve_regs(float2 (&x)[E], Mul mul, unsigned zero_waves = 0) {
    fwd<false>(x);
    rows_to_cols(x);
    if (wave_is_zero(zero_waves)) {
      zero_column_pass(x);
    } else {
      fwd<true>(x);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (E > 16) {
        // long lanes: eight points' multiplier loads at a time (register budget)
#pragma unroll
        for (int k = 0; k < E; ++k) {
          x[k] = mul(QL * k + q, rc, x[k]);
          if ((k & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
      } else {
#pragma unroll
        for (int k = 0; k < E; ++k) x[k] = mul(QL * k + q, rc, x[k]);
        // the multiplier's loads (a propagator: L2-resident) all issued before the products
        __builtin_amdgcn_sched_group_barrier(0x0020, 2 * E, 0);
        __builtin_amdgcn_sched_group_barrier(0x0002, 1024, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
      inv<true>(x);
    }
    cols_to_rows(x);
    inv<false>(x);
  }
  // in-register forward -> product with the scrambled-order plane mul (conjugated with
  // CONJ) -> inverse, the multiplier's loads issued before the forward transform so
  // their latency hides behind it
  // x[k] *= mul (conjugated with CONJ) at the column-phase points, the multiplier loaded
  // eight points at a time: for long lanes (E = 48 at n = 192) holding all E values
  // next to the E data points exceeds the register budget
  template <bool CONJ>
  __device__ __forceinline__ void mul_paced(float2 (&x)[E], const float2* mul) {
#pragma unroll
    for (int k0 = 0; k0 < E; k0 += 8) {
      float2 mv[8];
#pragma unroll
      for (int u = 0; u < 8 && k0 + u < E; ++u) mv[u] = mul[(QL * (k0 + u) + q) * N + rc];
#pragma unroll
      for (int u = 0; u < 8 && k0 + u < E; ++u) x[k0 + u] = CONJ ? cmulc(x[k0 + u], mv[u]) : cmul(x[k0 + u], mv[u]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // forward -> product with the scrambled-order plane mul -> inverse, mul loaded paced
  template <bool CONJ>
  __device__ __forceinline__ void convolve_paced(float2 (&x)[E], const float2* mul, unsigned zero_waves = 0) {
    fwd<false>(x);
    rows_to_cols(x);
    if (wave_is_zero(zero_waves)) {
      zero_column_pass(x);
    } else {
      fwd<true>(x);
      __builtin_amdgcn_sched_barrier(0);
      mul_paced<CONJ>(x, mul);
      inv<true>(x);
    }
    cols_to_rows(x);
    inv<false>(x);
  }
  template <bool CONJ>
  __device__ __forceinline__ void convolve_prefetched(float2 (&x)[E], const float2* mul, unsigned zero_waves = 0) {
    if constexpr (E > 16) {
      convolve_paced<CONJ>(x, mul, zero_waves);
      return;
    }
    const bool zero = wave_is_zero(zero_waves);
    float2 mv[E];
    if (!zero) {
#pragma unroll
      for (int k = 0; k < E; ++k) mv[k] = mul[(QL * k + q) * N + rc];
    }
    fwd<false>(x);
    rows_to_cols(x);
    if (zero) {
      zero_column_pass(x);
    } else {
      fwd<true>(x);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < E; ++k) x[k] = CONJ ? cmulc(x[k], mv[k]) : cmul(x[k], mv[k]);
      inv<true>(x);
    }
    cols_to_rows(x);
    inv<false>(x);
  }
  // as convolve_prefetched, the multiplier's loads issued after the row -> column
  // transpose instead: in flight during the forward column pass only
  template <bool CONJ>
  __device__ __forceinline__ void convolve_midfetch(float2 (&x)[E], const float2* mul, unsigned zero_waves = 0) {
    if constexpr (E > 16) {
      convolve_paced<CONJ>(x, mul, zero_waves);
      return;
    }
    fwd<false>(x);
    rows_to_cols(x);
    if (wave_is_zero(zero_waves)) {
      zero_column_pass(x);
    } else {
      float2 mv[E];
#pragma unroll
      for (int k = 0; k < E; ++k) mv[k] = mul[(QL * k + q) * N + rc];
      fwd<true>(x);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < E; ++k) x[k] = CONJ ? cmulc(x[k], mv[k]) : cmul(x[k], mv[k]);
      inv<true>(x);
    }
    cols_to_rows(x);
    inv<false>(x);
  }
  template <class Src, class Mul, class Dst, class Pin>
  __device__ __forceinline__ void convolve(Src src, Mul mul, Dst dst, Pin pin) {
    float2 x[E];
    load_rows(src, x, pin);
    convolve_regs(x, mul);
    store_rows(dst, x);
  }
  template <class Src, class Mul, class Dst>
  __device__ __forceinline__ void convolve(Src src, Mul mul, Dst dst) {
    convolve(src, mul, dst, [] {});
  }
};

// Load the twiddle table into LDS, arranged by step: entry QL k + q holds W_N^(q k)
// (q < QL, k < E: QL E = N entries), so lane q reads its E factors at one base plus
// compile-time offsets, and the QL lanes of a group read adjacent entries (no bank
// conflicts).  tw_glob holds W_N^e, e < N (host-computed in FP64).
template <int N>
__device__ __forceinline__ void load_twiddles(float2* tw_lds, const float2* tw_glob) {
  constexpr int QL = Plane<N>::QL;
  for (int i = threadIdx.x; i < N; i += Plane<N>::T) tw_lds[i] = tw_glob[(i % QL) * (i / QL)];
}

}  // namespace phx
