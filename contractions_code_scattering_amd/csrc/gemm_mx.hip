// Mixed-precision block-complex GEMM on the BF16 matrix cores (gfx950).
//
// Every real value x of an operand is held as three BF16 pieces, x = x0 + x1 + x2 exactly
// (each piece the round-to-nearest BF16 of what the previous ones leave; an FP32 significand
// has 24 bits = 3 x 8, so nothing is lost). A real product a*b is formed from the six piece
// products of weight >= 2^-16:
//   a0b0 + [a0b1 + a1b0 + a0b2 + a1b1 + a2b0]    (dropped terms are below 2^-24 |a||b|)
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, and the complex product from four real
// products (re = ArBr - AiBi, im = ArBi + AiBr): 24 MFMAs per 16x16x32 complex step. The leading
// a0b0 terms and the bracketed correction terms accumulate in separate registers, so the many
// small corrections never round the leading sum.
//
// Operands arrive already split ("split storage", lc_internal.h): the epilogue of the GEMM that
// produces a matrix writes its pieces once, and the inputs are split once by
// lc_split_convert_kernel. The main loop therefore only moves 16-byte BF16 vectors from global
// memory to LDS and feeds the matrix cores; view phases (powers of i) and conjugation are sign
// flips and a re/im plane exchange, applied on the way into LDS, or, where the chunk is staged by
// direct global -> LDS loads (Nv a multiple of 64), to the fragments read from LDS. Output is
// FP32 complex and/or
// split pieces; fused closing traces accumulate in FP64 with the same descriptors and per-tile
// partials as the FP64 kernel.
//
// Workgroup: 256 threads = 4 wave64 (2 x 2), 64 x 64 output tile of one spin block, K step 32.
// The MFMA is issued with the operands exchanged (B tile as the first operand), which yields the
// transposed product: lane l then holds four consecutive columns 4(l/16)..+3 of row l%16 of C,
// so the epilogue writes 8-byte piece vectors and 32-byte FP32 vectors. A workgroup processes
// `bpw` consecutive spin blocks of one product (bpw = 4 for Nv <= 64): prologue, elemental tile
// and trace reductions are amortized over four blocks, and the fetch pipeline runs across block
// boundaries.
#include "lc_device.h"

namespace lc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float fx4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ fx4v mfma_bf16(const bf16x8& a, const bf16x8& b, fx4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Result layout of v_mfma_f32_16x16x32_bf16: element q of lane l is row 4(l/16)+q, column l%16
// (measured with lc_mfma_bf16_selftest_kernel).
__device__ __forceinline__ int acc_row_bf16(int l, int q) { return ((l >> 4) << 2) + q; }

// Two FP32 -> packed round-to-nearest BF16 pair (one v_cvt_pk_bf16_f32), a in the low half.
__device__ __forceinline__ unsigned cvt2(float a, float b) {
  const bf16x2 v = __builtin_convertvector((f2v){a, b}, bf16x2);
  return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float lo_f(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float hi_f(unsigned u) { return __uint_as_float(u & 0xffff0000u); }

// Exact three-way split of a pair: p0, p1, p2 hold the pieces of (a, b), a in the low half.
__device__ __forceinline__ void split2(float a, float b, unsigned& p0, unsigned& p1, unsigned& p2) {
  p0 = cvt2(a, b);
  const float a1 = a - lo_f(p0), b1 = b - hi_f(p0);
  p1 = cvt2(a1, b1);
  const float a2 = a1 - lo_f(p1), b2 = b1 - hi_f(p1);
  p2 = cvt2(a2, b2);
}

constexpr int MX_TM = 64;
constexpr int MX_KT = 32;
// LDS line layout: a line holds its 32 K values as four 16-byte groups, group g stored at position
// g ^ ((line / 4) & 3). A fragment read (16 consecutive lines, one K group each) then touches 16
// distinct groups of four banks, and so does a 16-byte store of 4 lanes x 16 lines.
__device__ __forceinline__ int swz(int line) { return (line >> 2) & 3; }
constexpr unsigned SIGN2 = 0x80008000u;

// Position of element (line, k) in one Nv x Nv plane of split storage. For Nv a multiple of 64 the
// plane is kept in K panels of 32: panel k / 32 holds the 32 values of every line, line after
// line, so that the 64 x 32 chunk of a tile that the main loop stages is one contiguous 4 KB run
// per plane (whole cache lines). Otherwise lines are contiguous (row-major). A line is a row of a
// row-major operand or a column of a transposed one.
__device__ __forceinline__ long long plane_idx(int line, int k, int nv) {
  return (nv & 63) == 0 ? ((long long)(k >> 5) * nv + line) * 32 + (k & 31) : (long long)line * nv + k;
}

// Column of a 64-wide B tile held by LDS line `ph`. A wave's 16 x 16 MFMA block n reads the 16
// consecutive lines 32 wn + 16 n + i (i = 0..15, conflict-free with the LDS swizzle) and produces
// the columns those lines hold. Line 32 wn + 16 n + 4 a + b holds column 32 wn + 8 a + 4 n + b, so
// that the four columns a lane gets from block 0 and the four from block 1 are eight adjacent
// columns, and a lane stores each piece plane of a row as one 16-byte vector.
__device__ __forceinline__ int col_of_line(int ph) {
  return (ph & 32) + 8 * ((ph >> 2) & 3) + 4 * ((ph >> 4) & 1) + (ph & 3);
}

// Block (a,b) of a split view: base of its six planes, element flags, exact phase.
struct SplitRef {
  const unsigned short* p;
  int flags;
  float sr, si;
};

__device__ __forceinline__ SplitRef split_block(const LcView& v, long long off, int a, int b,
                                                long long nv2) {
  SplitRef r;
  long long e = off;
  r.flags = v.flags;
  r.sr = 1.f;
  r.si = 0.f;
  if (v.kind != LC_KRON) {
    const int ra = v.perm_r[a], cb = v.perm_c[b];
    const int blk = (v.flags & LC_TS) ? (cb * 4 + ra) : (ra * 4 + cb);
    e += (long long)blk * nv2;
    r.sr = (float)(v.ph_r[a][0] * v.ph_c[b][0] - v.ph_r[a][1] * v.ph_c[b][1]);
    r.si = (float)(v.ph_r[a][0] * v.ph_c[b][1] + v.ph_r[a][1] * v.ph_c[b][0]);
  }
  r.p = reinterpret_cast<const unsigned short*>(v.base) + 6 * e;
  return r;
}

// Six planes of 8 consecutive-K values of one line: element (line, k + e) of a plane sits at
// plane_idx(line, k + e) (A rows / B columns are contiguous along K, enforced by the planner; k is a
// multiple of 8, so the 8 values never straddle a K panel).
__device__ __forceinline__ void fetch_split(const unsigned short* p, int line, int k, int nv, long long nv2,
                                            bool vec, u32x4 (&v)[6]) {
  if (vec) {
#pragma unroll
    for (int q = 0; q < 6; q++)
      v[q] = gload<u32x4>(p + q * nv2 + plane_idx(line, k, nv));
  } else {
#pragma unroll
    for (int q = 0; q < 6; q++) {
      unsigned w[4];
#pragma unroll
      for (int h = 0; h < 4; h++) {
        unsigned lo = 0, hi = 0;
        if (line < nv && k + 2 * h < nv) lo = gload<unsigned short>(p + q * nv2 + plane_idx(line, k + 2 * h, nv));
        if (line < nv && k + 2 * h + 1 < nv)
          hi = gload<unsigned short>(p + q * nv2 + plane_idx(line, k + 2 * h + 1, nv));
        w[h] = lo | (hi << 16);
      }
      v[q] = u32x4{w[0], w[1], w[2], w[3]};
    }
  }
}

// FULL: every tile is complete (Nv a multiple of 64) and every operand and output is 16-byte
// aligned (checked by the launcher's caller); the kernel then has no element-wise fallback
// paths. Besides saving the bounds checks this matters for the schedule: the compiler's
// wait-count analysis merges all paths, so registers written by a fallback load would force a
// wait for the whole prefetch at the top of the matrix-core section.
//
// SW (sandwich, with FULL and Nv = 64): every group is C = alpha * (L (x) 1) M (R (x) 1) with L = A
// and R = B2 elementals and M = B a D-space view. Per output block (a,c) the workgroup first
// forms T = L M_ac (K = Nv), then, with T taken from the accumulators into LDS as the A operand
// (rounded to FP32 and split exactly, as a stored intermediate would be), C_ac = T R. T never
// leaves the chip. Its two K chunks are staged in two A buffers at once, so that no accumulator
// copy has to wait in registers for the second one. With a second right factor B3 the same T
// also forms T B3 (third term), which is only traced (Yx): one product L M_ac serves two
// sandwiches.
template <bool FULL, bool SW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
lc_gemm_mx_kernel(const LcGemmGroup* __restrict__ groups, int ng, int nv, long long total) {
  typedef float T;
  // planes [piece * 2 + re/im][line][k]
  __shared__ __attribute__((aligned(16))) __bf16 As[SW ? 2 : 1][6][MX_TM][MX_KT];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[6][MX_TM][MX_KT];
  __shared__ double red[8];

  const long long wg = xcd_order(total);
  if (wg < 0) return;
  const int gi = find_group(groups, ng, wg);
  const LcGemmGroup& g = groups[gi];
  const int tiles = g.tiles;
  const int tt = tiles * tiles;
  const int bpw = g.bpw;
  const long long local = wg - g.wg_begin;
  const long long per_item = 16LL * tt;            // output tiles per item
  const long long wg_per_item = per_item / bpw;    // workgroups per item
  const long long item = local / wg_per_item;
  const int r0 = (int)(local - item * wg_per_item) * bpw;  // first tile of this workgroup

  long long co[3];
  decode_item(item, g.n, co);
  const long long nv2 = (long long)nv * nv;
  const long long offA = voff(g.A.s, co), offB = voff(g.B.s, co);
  const bool akron = g.A.kind == LC_KRON, bkron = g.B.kind == LC_KRON;
  const bool vec_ok = FULL || (nv & 7) == 0;  // 16-byte plane vectors stay aligned

  // sandwich: term 0 is L M_ac, term 1 is T R, term 2 (if any) T R'
  const bool two_r = SW && g.B3.base != nullptr;
  const int nterm = SW ? (two_r ? 3 : 2) : (akron || bkron) ? 1 : 4;
  const int nk = (nv + MX_KT - 1) / MX_KT;
  const int cpb = nterm * nk;  // chunks per tile
  const int nchunk = cpb * bpw;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int s_line = tid >> 2, s_kg = tid & 3;  // staging: line and 8-wide K group of this thread
  const int s_pos = s_kg ^ swz(s_line);           // its position in the LDS line

  fx4v bre[2][2], bim[2][2], sre[2][2], sim[2][2];  // leading (a0b0) and correction sums
  auto zero_acc = [&]() {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        bre[i][j] = fx4v{0, 0, 0, 0}; bim[i][j] = fx4v{0, 0, 0, 0};
        sre[i][j] = fx4v{0, 0, 0, 0}; sim[i][j] = fx4v{0, 0, 0, 0};
      }
  };
  zero_acc();

  // Kernels with no bounds checks stage operand chunks through the 16-byte global -> LDS loads
  // (lane l of a wave writes LDS bytes 16 l .. 16 l + 15 past the wave's base), so no staging
  // registers and no LDS stores: the phase / conjugation masks are applied to the fragments
  // instead. The sandwich's T still goes from the accumulators to LDS by ordinary stores.
  constexpr bool DMA = FULL;
  // staging state of the chunk in flight
  u32x4 va[6], vb[6];
  int a_swap = 0;
  unsigned a_mr = 0, a_mi = 0, b_mi = 0;
  int a_t2 = 0, a_kc = 0;  // sandwich: term of the chunk (>= 1: T R or T R'), its K chunk
  int c_asel = 0;          // A buffer of the chunk in LDS

  auto tile_of = [&](int tl, int& a, int& c, int& tm, int& tn) {
    const int rr = r0 + tl;
    const int blk = rr / tt;
    const int rem = rr - blk * tt;
    tm = rem / tiles;
    tn = rem - tm * tiles;
    a = g.col_tiles ? (blk & 3) : (blk >> 2);
    c = g.col_tiles ? (blk >> 2) : (blk & 3);
  };

  // Chunk cursor of the fetch pipeline: tile, term (spin index b of the product) and K chunk of
  // the next chunk to fetch, advanced incrementally. Everything that depends only on the tile and
  // the term (block pointers of this thread's staging line, phase and conjugation masks) is
  // worked out once per term, not once per chunk.
  int cur_tl = 0, cur_term = 0, cur_kc = 0;
  int cur_tm = 0, cur_tn = 0;
  const unsigned short* pA = nullptr;  // this thread's staging line of the A block, K = 0
  const unsigned short* pB = nullptr;
  int t_swap = 0;
  unsigned t_mr = 0, t_mi = 0, t_bmi = 0;
  bool t_alA = true, t_alB = true;  // 16-byte alignment of the block bases (general kernel)

  // How the combined view phase (carried by A) and the conjugations apply: with A' = conj?(A)
  // the stored plane pair (R, I),
  //   phase  1: (R, I)      -1: (-R, -I)      i: (-I, R)      -i: (I, -R)
  // i.e. R goes to plane slot swap with sign mask mr, I to slot 1 - swap with mask mi.
  auto set_term = [&]() {
    int a, c;
    tile_of(cur_tl, a, c, cur_tm, cur_tn);
    if (SW && cur_term >= 1) {
      // T R: A comes from the accumulators, B is the right elemental (no phase)
      const LcView& R = cur_term == 1 ? g.B2 : g.B3;
      t_swap = 0;
      t_mr = 0u;
      t_mi = 0u;
      t_bmi = (R.flags & LC_CONJ) ? SIGN2 : 0u;
      pB = reinterpret_cast<const unsigned short*>(R.base) + 6 * voff(R.s, co) +
           plane_idx(cur_tn * MX_TM + col_of_line(s_line), (DMA ? s_pos : s_kg) * 8, nv);
      return;
    }
    int ab, bb, ba, bc;
    if (akron) { ab = a; bb = a; ba = a; bc = c; }
    else if (bkron) { ab = a; bb = c; ba = c; bc = c; }
    else { ab = a; bb = cur_term; ba = cur_term; bc = c; }
    const SplitRef A = split_block(g.A, offA, ab, bb, nv2);
    const SplitRef B = split_block(g.B, offB, ba, bc, nv2);
    const float sr = A.sr * B.sr - A.si * B.si, si = A.sr * B.si + A.si * B.sr;
    const unsigned ca = (A.flags & LC_CONJ) ? SIGN2 : 0u;
    if (si == 0.f) {
      t_swap = 0;
      t_mr = sr < 0.f ? SIGN2 : 0u;
      t_mi = t_mr ^ ca;
    } else {
      t_swap = 1;
      // phase i: slot re <- -I', slot im <- R;  phase -i: slot re <- I', slot im <- -R
      t_mr = si > 0.f ? 0u : SIGN2;         // sign applied to R (lands in the im slot)
      t_mi = (si > 0.f ? SIGN2 : 0u) ^ ca;  // sign applied to I (lands in the re slot)
    }
    t_bmi = (B.flags & LC_CONJ) ? SIGN2 : 0u;
    t_alA = (reinterpret_cast<uintptr_t>(A.p) & 15) == 0;
    t_alB = (reinterpret_cast<uintptr_t>(B.p) & 15) == 0;
    // DMA: the lane's 16 bytes land at LDS position s_kg of its line, which holds K group
    // s_kg ^ swz(line) = s_pos
    const int kg = DMA ? s_pos : s_kg;
    pA = A.p + plane_idx(cur_tm * MX_TM + s_line, kg * 8, nv);
    pB = B.p + plane_idx(cur_tn * MX_TM + col_of_line(s_line), kg * 8, nv);
  };
  set_term();

  // Fetch the chunk under the cursor into registers (with its masks), then advance the cursor.
  auto fetch = [&]() {
    const int k0 = cur_kc * MX_KT;
    a_swap = t_swap;
    a_mr = t_mr;
    a_mi = t_mi;
    b_mi = t_bmi;
    a_t2 = SW ? cur_term : 0;
    a_kc = cur_kc;
    // chunk kc starts k0 values further along each line: in the K-panel layout (Nv a multiple of
    // 64) that is k0 / 32 panels of 32 Nv values further
    const long long kofs = (nv & 63) == 0 ? (long long)k0 * nv : k0;
    if (DMA) {
#pragma unroll
      for (int q = 0; q < 6; q++) {
        if (!SW || !a_t2)
          __builtin_amdgcn_global_load_lds((const void*)(pA + q * nv2 + kofs),
                                           (__attribute__((address_space(3))) void*)&As[0][q][wave * 16][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(pB + q * nv2 + kofs),
                                         (__attribute__((address_space(3))) void*)&Bs[q][wave * 16][0], 16, 0, 0);
      }
    } else {
      // general kernel (SW and FULL kernels always take the branch above)
      const bool fullA = vec_ok && t_alA && (cur_tm + 1) * MX_TM <= nv && k0 + MX_KT <= nv;
      const bool fullB = vec_ok && t_alB && (cur_tn + 1) * MX_TM <= nv && k0 + MX_KT <= nv;
      const int k = k0 + s_kg * 8;
      fetch_split(pA - plane_idx(cur_tm * MX_TM + s_line, s_kg * 8, nv), cur_tm * MX_TM + s_line, k, nv, nv2, fullA,
                  va);
      fetch_split(pB - plane_idx(cur_tn * MX_TM + col_of_line(s_line), s_kg * 8, nv),
                  cur_tn * MX_TM + col_of_line(s_line), k, nv, nv2, fullB, vb);
    }
    if (++cur_kc == nk) {
      cur_kc = 0;
      if (++cur_term == nterm) {
        cur_term = 0;
        ++cur_tl;
      }
      if (cur_tl < bpw) set_term();
    }
  };

  auto store = [&]() {
    c_asel = 0;
    if (DMA && !SW) return;
    if (SW && a_t2) {
      c_asel = a_kc;
      if (a_t2 == 1 && a_kc == 0) {
        // T is complete: this wave's 32 x 32 part goes to A buffer wn (T's columns are the K of
        // T R, 32 per chunk), split, in the layout of a staged A chunk
#pragma unroll
        for (int m = 0; m < 2; m++) {
          // the lane's eight adjacent columns of T (blocks n = 0, 1) are one 8-wide K group
          unsigned r01[2][3], r23[2][3], i01[2][3], i23[2][3];
#pragma unroll
          for (int n = 0; n < 2; n++) {
            const fx4v xr = bre[m][n] + sre[m][n], xi = bim[m][n] + sim[m][n];
            split2(xr[0], xr[1], r01[n][0], r01[n][1], r01[n][2]);
            split2(xr[2], xr[3], r23[n][0], r23[n][1], r23[n][2]);
            split2(xi[0], xi[1], i01[n][0], i01[n][1], i01[n][2]);
            split2(xi[2], xi[3], i23[n][0], i23[n][1], i23[n][2]);
          }
          const int line = wm * 32 + m * 16 + (lane & 15);
          const int e = ((lane >> 4) ^ swz(line)) * 8;
#pragma unroll
          for (int p = 0; p < 3; p++) {
            *reinterpret_cast<u32x4*>(&As[wn][2 * p][line][e]) = u32x4{r01[0][p], r23[0][p], r01[1][p], r23[1][p]};
            *reinterpret_cast<u32x4*>(&As[wn][2 * p + 1][line][e]) = u32x4{i01[0][p], i23[0][p], i01[1][p], i23[1][p]};
          }
        }
        zero_acc();
      }
      return;
    }
    if (DMA) return;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      *reinterpret_cast<u32x4*>(&As[0][2 * p + a_swap][s_line][s_pos * 8]) = va[2 * p] ^ a_mr;
      *reinterpret_cast<u32x4*>(&As[0][2 * p + 1 - a_swap][s_line][s_pos * 8]) = va[2 * p + 1] ^ a_mi;
      *reinterpret_cast<u32x4*>(&Bs[2 * p][s_line][s_pos * 8]) = vb[2 * p];
      *reinterpret_cast<u32x4*>(&Bs[2 * p + 1][s_line][s_pos * 8]) = vb[2 * p + 1] ^ b_mi;
    }
  };

  // per-tile epilogue: stores and fused trace accumulation (FP64, carried across tiles)
  double tsr[LC_MAX_FUSED_TRACES], tsi[LC_MAX_FUSED_TRACES];
#pragma unroll
  for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) { tsr[t] = 0; tsi[t] = 0; }
  double txr = 0, txi = 0;  // trace of the second sandwich product
  const T alr = T(g.alpha[0]), ali = T(g.alpha[1]);
  const bool vec_c = FULL || (nv & 3) == 0;

  // Per-tile epilogue: stores, and the fused trace pieces sum_ij C_ij Y_ji in FP64. Every load
  // of the trace operands (all traces, all 16 values of this lane) is issued before the first
  // one is used, so the tile pays one memory round trip instead of one per value.
  // second: the tile of T R' (sandwich term 2): no stores, one trace (Yx)
  auto epilogue = [&](int tl, bool second) {
    // keep the scheduler from hoisting the epilogue's loads into the matrix-core section
    __builtin_amdgcn_sched_barrier(0);
    int a, c, tm, tn;
    tile_of(tl, a, c, tm, tn);
    const long long cofs = voff(g.cs, co) + (long long)(a * 4 + c) * nv2;
    const int rbase = tm * MX_TM + wm * 32 + (lane & 15);
    // this lane's columns: cbase + 4 n + q for block n (eight adjacent columns, see col_of_line)
    const int cbase = tn * MX_TM + wn * 32 + 8 * (lane >> 4);
    // the tile's values (alpha applied); the accumulators are free from here on
    float zr_[2][2][4], zi_[2][2][4];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n = 0; n < 2; n++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const T xr = bre[m][n][q] + sre[m][n][q], xi = bim[m][n][q] + sim[m][n][q];
          zr_[m][n][q] = second ? xr : alr * xr - ali * xi;
          zi_[m][n][q] = second ? xi : alr * xi + ali * xr;
        }
    zero_acc();
    // trace operand of trace t: this lane's 16 values of f(Y), element (col, row), loaded with no
    // use in between; block phase and conjugation kept aside
    struct YTile {
      f2v v[2][2][4];
      double pr, pi;
      float cj;
    };
    auto load_y = [&](int t, YTile& y) {
      y.pr = 0; y.pi = 0; y.cj = 1.f;
      BlockRef<T> Y;
      Y.zero = true;
      if (second) {
        if (t == 0) Y = view_block<T>(g.Yx, voff(g.Yx.s, co), c, a, nv2);
      } else if (t < g.ntr) {
        Y = view_block<T>(g.Y[t], voff(g.Y[t].s, co), c, a, nv2);
      }
      if (!Y.zero) { y.pr = Y.sr; y.pi = Y.si; y.cj = (Y.flags & LC_CONJ) ? -1.f : 1.f; }
      const bool tv = (Y.flags & LC_TV) != 0;
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) {
          const int row = rbase + m * 16, col = cbase + 4 * n;
          if (FULL && tv) {
            // f(Y)(col + q, row) = Y(row, col + q): four consecutive complex values
            const float* p = reinterpret_cast<const float*>(Y.p + (long long)row * nv + col);
            const fx4v v01 = Y.zero ? fx4v{0.f, 0.f, 0.f, 0.f} : gload<fx4v>(p);
            const fx4v v23 = Y.zero ? fx4v{0.f, 0.f, 0.f, 0.f} : gload<fx4v>(p + 4);
            y.v[m][n][0] = f2v{v01[0], v01[1]};
            y.v[m][n][1] = f2v{v01[2], v01[3]};
            y.v[m][n][2] = f2v{v23[0], v23[1]};
            y.v[m][n][3] = f2v{v23[2], v23[3]};
          } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
              const long long idx = tv ? (long long)row * nv + col + q : (long long)(col + q) * nv + row;
              y.v[m][n][q] = (!Y.zero && (FULL || (row < nv && col + q < nv))) ? gload<f2v>(Y.p + idx)
                                                                               : f2v{0.f, 0.f};
            }
          }
        }
    };
    // every loaded value is used on every path (traces past ntr sum zeros): a load left pending
    // on some path would make the compiler wait for all memory traffic at the next reuse of its
    // register, i.e. for the whole prefetch at the top of the matrix-core section
    YTile ycur, ynext;
    load_y(0, ycur);
#pragma unroll
    for (int m = 0; m < 2; m++) {
      // full tiles: the split pieces of both blocks, stored together below (16 bytes per plane)
      unsigned mr01[2][3], mr23[2][3], mi01[2][3], mi23[2][3];
      const bool merged = FULL && !second && g.Cs && !g.cs_trans;
#pragma unroll
      for (int n = 0; n < 2; n++) {
        // this lane's values: row `row`, columns col..col+3
        const int row = rbase + m * 16;
        const int col = cbase + 4 * n;
        const float* zr = zr_[m][n];
        const float* zi = zi_[m][n];
        const bool rin = FULL || row < nv;
        const bool full4 = FULL || (rin && col + 3 < nv);
        if (second) continue;
        if (g.C) {
          float* C = reinterpret_cast<float*>(g.C) + 2 * (cofs + (long long)row * nv + col);
          if (FULL || (full4 && vec_c && (cofs & 1) == 0)) {
            gstore(C, fx4v{zr[0], zi[0], zr[1], zi[1]});
            gstore(C + 4, fx4v{zr[2], zi[2], zr[3], zi[3]});
          } else if (rin) {
#pragma unroll
            for (int q = 0; q < 4; q++)
              if (col + q < nv) gstore(C + 2 * q, f2v{zr[q], zi[q]});
          }
        }
        if (g.Cs) {
          unsigned r01[3], r23[3], i01[3], i23[3];
          split2(zr[0], zr[1], r01[0], r01[1], r01[2]);
          split2(zr[2], zr[3], r23[0], r23[1], r23[2]);
          split2(zi[0], zi[1], i01[0], i01[1], i01[2]);
          split2(zi[2], zi[3], i23[0], i23[1], i23[2]);
          unsigned short* S = reinterpret_cast<unsigned short*>(g.Cs) + 6 * cofs;
          if (merged) {
#pragma unroll
            for (int p = 0; p < 3; p++) {
              mr01[n][p] = r01[p];
              mr23[n][p] = r23[p];
              mi01[n][p] = i01[p];
              mi23[n][p] = i23[p];
            }
          } else if (!g.cs_trans && (FULL || (full4 && vec_c && (cofs & 1) == 0))) {
            const long long e = plane_idx(row, col, nv);  // col % 4 == 0: four values in one K panel
#pragma unroll
            for (int p = 0; p < 3; p++) {
              gstore(S + (2 * p) * nv2 + e, u32x2{r01[p], r23[p]});
              gstore(S + (2 * p + 1) * nv2 + e, u32x2{i01[p], i23[p]});
            }
          } else if (rin) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
              if (col + q >= nv) continue;
              const long long e = g.cs_trans ? plane_idx(col + q, row, nv) : plane_idx(row, col + q, nv);
              const int sh = (q & 1) * 16;
#pragma unroll
              for (int p = 0; p < 3; p++) {
                const unsigned wr = q < 2 ? r01[p] : r23[p], wi = q < 2 ? i01[p] : i23[p];
                gstore(S + (2 * p) * nv2 + e, (unsigned short)(wr >> sh));
                gstore(S + (2 * p + 1) * nv2 + e, (unsigned short)(wi >> sh));
              }
            }
          }
        }
      }
      if (merged) {
        // columns cbase..cbase+7 of row rbase + 16 m: one K panel, 16-byte aligned
        unsigned short* S = reinterpret_cast<unsigned short*>(g.Cs) + 6 * cofs + plane_idx(rbase + m * 16, cbase, nv);
#pragma unroll
        for (int p = 0; p < 3; p++) {
          gstore(S + (2 * p) * nv2, u32x4{mr01[0][p], mr23[0][p], mr01[1][p], mr23[1][p]});
          gstore(S + (2 * p + 1) * nv2, u32x4{mi01[0][p], mi23[0][p], mi01[1][p], mi23[1][p]});
        }
      }
    }
    // traces, the loads of trace t+1 in flight while trace t is summed
#pragma unroll
    for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) {
      if (t + 1 < LC_MAX_FUSED_TRACES) load_y(t + 1, ynext);
      // four-term FP32 partial sums (rounding far below that of the FP32 accumulation of the
      // product itself), carried on in FP64
      double sr = 0, si = 0;
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) {
          float fr = 0.f, fi = 0.f;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const float yr = ycur.v[m][n][q][0], yi = ycur.cj * ycur.v[m][n][q][1];
            const float xr = zr_[m][n][q], xi = zi_[m][n][q];
            fr += xr * yr - xi * yi;
            fi += xr * yi + xi * yr;
          }
          sr += fr;
          si += fi;
        }
      // the Y block phase is uniform over the tile
      const double dr = sr * ycur.pr - si * ycur.pi, di = sr * ycur.pi + si * ycur.pr;
      if (!second) {
        tsr[t] += dr;
        tsi[t] += di;
      } else if (t == 0) {
        txr += dr;
        txi += di;
      }
      if (t + 1 < LC_MAX_FUSED_TRACES) ycur = ynext;
    }
  };

  fetch();
  store();
  __syncthreads();
  // fragment lines of this lane and the LDS offsets of its K group in them
  int al[2], ao[2], bl[2], bo[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    al[i] = wm * 32 + i * 16 + (lane & 15);
    bl[i] = wn * 32 + i * 16 + (lane & 15);
    ao[i] = ((lane >> 4) ^ swz(al[i])) * 8;
    bo[i] = ((lane >> 4) ^ swz(bl[i])) * 8;
  }
  if (DMA) {
    __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0): this wave's LDS DMA has landed
    __syncthreads();
  }
  for (int ch = 0; DMA && ch < nchunk; ch++) {
    const bool more = ch + 1 < nchunk;
    // all fragments of the chunk, with its phase / conjugation masks, before the buffer is refilled
    bf16x8 bf[2][3][2], af[2][3][2];  // [n or m][piece][re/im]
    const int pre = a_swap;
    const unsigned m0 = pre ? a_mi : a_mr, m1 = pre ? a_mr : a_mi;
#pragma unroll
    for (int n = 0; n < 2; n++)
#pragma unroll
      for (int p = 0; p < 3; p++) {
        bf[n][p][0] = *reinterpret_cast<const bf16x8*>(&Bs[2 * p][bl[n]][bo[n]]);
        bf[n][p][1] = __builtin_bit_cast(
            bf16x8, __builtin_bit_cast(u32x4, *reinterpret_cast<const bf16x8*>(&Bs[2 * p + 1][bl[n]][bo[n]])) ^ b_mi);
      }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int p = 0; p < 3; p++) {
        af[m][p][0] = __builtin_bit_cast(
            bf16x8,
            __builtin_bit_cast(u32x4, *reinterpret_cast<const bf16x8*>(&As[SW ? c_asel : 0][2 * p + pre][al[m]][ao[m]])) ^ m0);
        af[m][p][1] = __builtin_bit_cast(
            bf16x8,
            __builtin_bit_cast(u32x4, *reinterpret_cast<const bf16x8*>(&As[SW ? c_asel : 0][2 * p + 1 - pre][al[m]][ao[m]])) ^ m1);
      }
    __syncthreads();  // every wave holds its fragments: the buffer may be refilled
    if (more) fetch();
#pragma unroll
    for (int m = 0; m < 2; m++) {
      bf16x8 nai[3];
#pragma unroll
      for (int p = 0; p < 3; p++)
        nai[p] = __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, af[m][p][1]) ^ SIGN2);
#pragma unroll
      for (int n = 0; n < 2; n++) {
        const int PI[5] = {2, 1, 0, 1, 0};
        const int PJ[5] = {0, 1, 2, 0, 1};
#pragma unroll
        for (int s = 0; s < 5; s++) {
          const int i = PI[s], j = PJ[s];
          sre[m][n] = mfma_bf16(bf[n][j][0], af[m][i][0], sre[m][n]);
          sre[m][n] = mfma_bf16(bf[n][j][1], nai[i], sre[m][n]);
          sim[m][n] = mfma_bf16(bf[n][j][1], af[m][i][0], sim[m][n]);
          sim[m][n] = mfma_bf16(bf[n][j][0], af[m][i][1], sim[m][n]);
        }
        bre[m][n] = mfma_bf16(bf[n][0][0], af[m][0][0], bre[m][n]);
        bre[m][n] = mfma_bf16(bf[n][0][1], nai[0], bre[m][n]);
        bim[m][n] = mfma_bf16(bf[n][0][1], af[m][0][0], bim[m][n]);
        bim[m][n] = mfma_bf16(bf[n][0][0], af[m][0][1], bim[m][n]);
      }
    }
    {
      const int tl = ch / cpb, kin = ch - tl * cpb;
      if (!SW) {
        if (kin == cpb - 1) epilogue(tl, false);
      } else if (kin == 2 * nk - 1) {
        epilogue(tl, false);
      } else if (kin == 3 * nk - 1) {
        epilogue(tl, true);
      }
    }
    if (more) {
      store();  // sandwich: T of the finished term 0 into the A buffers (no global traffic)
      __builtin_amdgcn_s_waitcnt(0xF70);
      __syncthreads();
    }
  }
  for (int ch = 0; !DMA && ch < nchunk; ch++) {
    const bool more = ch + 1 < nchunk;
    if (more) fetch();
    bf16x8 bf[2][3][2];  // [n][piece][re/im]
#pragma unroll
    for (int n = 0; n < 2; n++)
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int ri = 0; ri < 2; ri++)
          bf[n][p][ri] = *reinterpret_cast<const bf16x8*>(&Bs[2 * p + ri][bl[n]][bo[n]]);
#pragma unroll
    for (int m = 0; m < 2; m++) {
      bf16x8 af[3][2], nai[3];
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int ri = 0; ri < 2; ri++)
          af[p][ri] = *reinterpret_cast<const bf16x8*>(&As[0][2 * p + ri][al[m]][ao[m]]);
#pragma unroll
      for (int p = 0; p < 3; p++)
        nai[p] = __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, af[p][1]) ^ SIGN2);
#pragma unroll
      for (int n = 0; n < 2; n++) {
        // correction terms (weights 2^-8 and 2^-16), then the leading term; B tile first, so
        // the accumulators hold the transposed product
        const int PI[5] = {2, 1, 0, 1, 0};
        const int PJ[5] = {0, 1, 2, 0, 1};
#pragma unroll
        for (int s = 0; s < 5; s++) {
          const int i = PI[s], j = PJ[s];
          sre[m][n] = mfma_bf16(bf[n][j][0], af[i][0], sre[m][n]);
          sre[m][n] = mfma_bf16(bf[n][j][1], nai[i], sre[m][n]);
          sim[m][n] = mfma_bf16(bf[n][j][1], af[i][0], sim[m][n]);
          sim[m][n] = mfma_bf16(bf[n][j][0], af[i][1], sim[m][n]);
        }
        bre[m][n] = mfma_bf16(bf[n][0][0], af[0][0], bre[m][n]);
        bre[m][n] = mfma_bf16(bf[n][0][1], nai[0], bre[m][n]);
        bim[m][n] = mfma_bf16(bf[n][0][1], af[0][0], bim[m][n]);
        bim[m][n] = mfma_bf16(bf[n][0][0], af[0][1], bim[m][n]);
      }
    }
    {
      const int tl = ch / cpb, kin = ch - tl * cpb;
      if (kin == cpb - 1) epilogue(tl, false);
    }
    __syncthreads();
    if (more) {
      store();
      __syncthreads();
    }
  }

#pragma unroll
  for (int t = 0; t < LC_MAX_FUSED_TRACES; t++) {
    if (t >= g.ntr) break;
    double sr = tsr[t], si = tsi[t];
    block_sum2(sr, si, red);
    if (tid == 0) {
      const double ar = g.tr_alpha[t][0], ai = g.tr_alpha[t][1];
      double* out = g.part[t] + 2 * (item * wg_per_item + r0 / bpw);
      typedef double d2v __attribute__((ext_vector_type(2)));
      gstore(out, d2v{ar * sr - ai * si, ar * si + ai * sr});
    }
  }
  if (two_r) {
    double sr = txr, si = txi;
    block_sum2(sr, si, red);
    if (tid == 0) {
      const double ar = g.trx_alpha[0], ai = g.trx_alpha[1];
      double* out = g.partx + 2 * (item * wg_per_item + r0 / bpw);
      typedef double d2v __attribute__((ext_vector_type(2)));
      gstore(out, d2v{ar * sr - ai * si, ar * si + ai * sr});
    }
  }
}

// Layout self-test of v_mfma_f32_16x16x32_bf16 with A(16x32), B(32x16) given as FP32 values that
// are exact in BF16: lane l feeds A[l%16][8(l/16)..+7] and B[8(l/16)..+7][l%16]. d[0..255] = D
// row-major through acc_row_bf16, d[256..511] = raw accumulator q of lane l at 256 + 4l + q.
__global__ void lc_mfma_bf16_selftest_kernel(const float* a, const float* b, float* d) {
  const int l = threadIdx.x;
  bf16x8 av, bv;
  for (int e = 0; e < 8; e++) {
    av[e] = (__bf16)a[(l & 15) * 32 + (l >> 4) * 8 + e];
    bv[e] = (__bf16)b[((l >> 4) * 8 + e) * 16 + (l & 15)];
  }
  fx4v c = fx4v{0, 0, 0, 0};
  c = mfma_bf16(av, bv, c);
  for (int q = 0; q < 4; q++) {
    d[acc_row_bf16(l, q) * 16 + (l & 15)] = c[q];
    d[256 + l * 4 + q] = c[q];
  }
}

// ------------------------------------------------------------------------------------------
// Inputs of the mixed path: per-slice Frobenius scale ||x_i||_F (FP64), then x_i / ||x_i||_F
// rounded to FP32 and split into the six BF16 planes, written twice: row-major planes (y_rows)
// and with every Nv x Nv block transposed (y_trans), so that each operand role finds its K
// dimension contiguous; optionally also the FP32 values themselves (y32, closing-trace factors).
// A slice is `len` complex values made of len / Nv^2 blocks.
__global__ void __launch_bounds__(256) lc_slice_norm_kernel(const double2* __restrict__ x,
                                                            long long len, long long stride,
                                                            double* __restrict__ scale) {
  __shared__ double red[8];
  const double2* p = x + (long long)blockIdx.x * stride;
  double s = 0, dummy = 0;
  for (long long e = threadIdx.x; e < len; e += 256) {
    const double2 v = p[e];
    s += v.x * v.x + v.y * v.y;
  }
  block_sum2(s, dummy, red);
  if (threadIdx.x == 0) scale[blockIdx.x] = s > 0 ? sqrt(s) : 1.0;
}

// One workgroup per 32 x 32 tile of one Nv x Nv block: the values are read and the row-major
// planes written along rows; the pieces go through LDS so that the transposed planes are written
// along rows of the transpose as well.
constexpr int SC_T = 32;
__global__ void __launch_bounds__(256) lc_split_convert_kernel(const double2* __restrict__ x,
                                                               unsigned short* __restrict__ yr,
                                                               unsigned short* __restrict__ yt,
                                                               float2* __restrict__ y32, long long len,
                                                               int nv, const double* __restrict__ scale) {
  __shared__ unsigned short pc[6][SC_T][SC_T + 1];
  const int tpb = (nv + SC_T - 1) / SC_T;  // tiles per block side
  const long long tile = blockIdx.x;
  const long long bi = tile / (tpb * tpb);  // global block index (slice * blocks per slice + block)
  const int tr = (int)((tile / tpb) % tpb), tc = (int)(tile % tpb);
  const long long nv2 = (long long)nv * nv;
  const double inv = 1.0 / scale[(bi * nv2) / len];
  const double2* xb = x + bi * nv2;
  unsigned short* br = yr + 6 * bi * nv2;
  unsigned short* bt = yt + 6 * bi * nv2;
  const int c = threadIdx.x & (SC_T - 1), r0 = threadIdx.x >> 5;  // 8 rows per pass
#pragma unroll
  for (int pass = 0; pass < SC_T / 8; pass++) {
    const int r = r0 + pass * 8;
    const int row = tr * SC_T + r, col = tc * SC_T + c;
    if (row < nv && col < nv) {
      const long long e = (long long)row * nv + col;  // FP64 input and FP32 copy: row-major
      const double2 v = xb[e];
      const float fr = (float)(v.x * inv), fi = (float)(v.y * inv);
      if (y32) y32[bi * nv2 + e] = make_float2(fr, fi);
      unsigned p0, p1, p2;
      split2(fr, fi, p0, p1, p2);
      const unsigned q[3] = {p0, p1, p2};
      const long long es = plane_idx(row, col, nv);
#pragma unroll
      for (int p = 0; p < 3; p++) {
        br[(2 * p) * nv2 + es] = (unsigned short)(q[p] & 0xffff);
        br[(2 * p + 1) * nv2 + es] = (unsigned short)(q[p] >> 16);
        pc[2 * p][r][c] = (unsigned short)(q[p] & 0xffff);
        pc[2 * p + 1][r][c] = (unsigned short)(q[p] >> 16);
      }
    }
  }
  __syncthreads();
  // transposed planes: element (row, col) is element (line col, k row); here thread c runs along row
#pragma unroll
  for (int pass = 0; pass < SC_T / 8; pass++) {
    const int r = r0 + pass * 8;  // a column of the tile
    const int row = tr * SC_T + c, col = tc * SC_T + r;
    if (row < nv && col < nv) {
      const long long et = plane_idx(col, row, nv);
#pragma unroll
      for (int q = 0; q < 6; q++) bt[q * nv2 + et] = pc[q][c][r];
    }
  }
}

}  // namespace lc

// variant 2: sandwich kernel (full tiles, Nv = 64); 1: full-tile kernel; 0: general kernel. The
// planner decides which conditions every group of the launch meets.
extern "C" hipError_t lc_launch_gemm_mx(int variant, const LcGemmGroup* d_groups, int ng, long long total_wg,
                                        int nv, hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  const long long per = (total_wg + 7) / 8;
  const dim3 grid((unsigned)(8 * per)), block(256);
  if (variant == 2)
    hipLaunchKernelGGL((lc::lc_gemm_mx_kernel<true, true>), grid, block, 0, s, d_groups, ng, nv, total_wg);
  else if (variant == 1)
    hipLaunchKernelGGL((lc::lc_gemm_mx_kernel<true, false>), grid, block, 0, s, d_groups, ng, nv, total_wg);
  else
    hipLaunchKernelGGL((lc::lc_gemm_mx_kernel<false, false>), grid, block, 0, s, d_groups, ng, nv, total_wg);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_mfma_bf16_selftest(const float* a, const float* b, float* d,
                                                   hipStream_t s) {
  hipLaunchKernelGGL(lc::lc_mfma_bf16_selftest_kernel, dim3(1), dim3(64), 0, s, a, b, d);
  return hipGetLastError();
}

// x: nslice slices of len complex doubles (contiguous, len a multiple of nv^2); y_rows, y_trans:
// the split planes (6 * len * nslice BF16 each); scale: nslice doubles.
extern "C" hipError_t lc_launch_split_rescale(const void* x, void* y_rows, void* y_trans, void* y32,
                                              long long len, int nv, long long nslice, double* scale,
                                              hipStream_t s) {
  if (nslice <= 0 || len <= 0) return hipSuccess;
  hipLaunchKernelGGL(lc::lc_slice_norm_kernel, dim3((unsigned)nslice), dim3(256), 0, s,
                     (const double2*)x, len, len, scale);
  const long long nv2 = (long long)nv * nv;
  const long long tpb = (nv + lc::SC_T - 1) / lc::SC_T;
  const long long tiles = (len * nslice / nv2) * tpb * tpb;
  hipLaunchKernelGGL(lc::lc_split_convert_kernel, dim3((unsigned)tiles), dim3(256), 0, s, (const double2*)x,
                     (unsigned short*)y_rows, (unsigned short*)y_trans, (float2*)y32, len, nv,
                     (const double*)scale);
  return hipGetLastError();
}
