// Exact-accumulation block-complex GEMM on the int8 matrix cores (gfx950), and the conversions
// into and out of digit storage.
//
// Number format (lc_internal.h). Every real value x of a stored line (a row of an A operand, a
// column of a B operand) is four balanced base-255 digits under the line's power-of-two exponent E
// and the fixed factor SIGMA = 63/32:
//   x = v / (SIGMA 2^(30-E)),   v = D0 255^3 + D1 255^2 + D2 255 + D3,   v = rint(x SIGMA 2^(30-E)),
// with D0..D3 in [-127, 127] (|v| < SIGMA 2^30 <= (255^4 - 1) / 2) and E the smallest integer such
// that every |x| of the line is below 2^E: truncation <= 2^(E-31) / SIGMA. SIGMA lets the leading
// digit use the whole int8 range: the largest products of a line then dominate the dropped digit
// orders by (127 / 65)^2 more than under a plain power of two. Base 255 instead of 256 keeps every
// digit's negation inside int8 (the products need negated digits for the complex product and for
// phases and conjugation).
//
// Product. For C = A B every digit pair (a, b) of total order o = a + b <= 3 is one
// v_mfma_i32_32x32x32_i8 per real product (10 pairs); the pairs of one order share an int32
// accumulator, |acc_o| <= 2 (o + 1) K 127^2 < 2^27 for K <= 512, so the integer sums are exact and
// the only rounding of a product is the digit truncation of its operands and the dropped orders
// 4..6, all near 2^-32 of the largest products (a numerical model of the format at D = 256 gives
// about 6e-9 of max |C| for random operands). The epilogue forms, per element,
//   S = ((acc_0 255 + acc_1) 255 + acc_2) 255 + acc_3          (exact in FP64, < 2^51)
//   C_ij = alpha * 2^(E_A,i + E_B,j - 60) * rn(S UNSCALE),   UNSCALE = rn(255^3 / SIGMA^2)
// (the one rounded step, rn, is relative 2^-53)
// and stores it in FP64, traces it in FP64 against FP64 closing factors, and/or writes it back as
// digits with a fresh exponent per D-row (the largest |C_ij| of the row over all four blocks,
// which the workgroup strip covers). The complex product takes four real products (re += Ar Br -
// Ai Bi, im += Ar Bi + Ai Br): 40 MFMAs per 32 x 32 output tile and K step of 32.
//
// Workgroup: NW wave64s own a strip of 32 rows x 32 NW columns of one item; wave w the 32 x 32
// tile at columns 32 w. Operands move in chunks of 32 lines x 32 K x 8 digit planes, one
// contiguous kilobyte per plane in memory, laid out in the order of the MFMA fragments (lane l:
// line l % 32, K bytes 16 (l / 32) .. +15). The A chunk of a step, shared by the waves, is loaded
// one step ahead into registers and written to a double-buffered LDS stage; every wave loads its
// own B fragments straight from global memory into registers (one coalesced 1 KB load per plane),
// each digit's planes reloaded for the next step right after their last use in the current one.
// One barrier per step. View phases (powers of i) and conjugation are applied to the fragments as
// digit negations and a choice of planes.
//
// The MFMA is issued with the B fragment first, so lane l holds row l % 32 of C and the 16
// columns 8 g + 4 (l / 32) + r (g, r < 4). The epilogue traces and converts from these registers
// (closing factors are laid out like them; digit rows are exchanged between the lane halves) and
// stages the tile in LDS only for FP64 outputs and FP64 closing factors, stored as whole rows.
#include <climits>
#include <type_traits>

#include "lc_device.h"

namespace lc {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef double d2v __attribute__((ext_vector_type(2)));
constexpr int NORD = 4;  // digit-pair orders 0..3 accumulated

__device__ __forceinline__ i32x16 mfma_i8(const i32x4& x, const i32x4& y, const i32x16& c) {
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(x, y, c, 0, 0, 0);
}

// Negation of four packed int8 digits in [-127, 127]: in the sign-flipped (biased) domain the
// digits are 1..255, whose negation 256 - b is ~b + 1 with no carry between bytes.
__device__ __forceinline__ unsigned neg4(unsigned x) { return ((x ^ 0x7F7F7F7Fu) + 0x01010101u) ^ 0x80808080u; }
__device__ __forceinline__ void neg16(i32x4& v) {
  v = i32x4{(int)neg4((unsigned)v[0]), (int)neg4((unsigned)v[1]), (int)neg4((unsigned)v[2]), (int)neg4((unsigned)v[3])};
}

// Exponent E of a nonzero value: |x| < 2^E (the frexp exponent); INT_MIN for zero, so that zeros
// never raise a line's exponent.
__device__ __forceinline__ int exp_of(double x) { return x == 0.0 ? INT_MIN : __builtin_amdgcn_frexp_exp(x); }

// Round-to-integer constant 1.5 2^52: fma(y, c, RND) rounds the exact y c to an integer Q (|Q| < 2^51)
// and holds Q mod 2^32 in its low word.
constexpr double RND = 6755399441055744.0;
__device__ __forceinline__ int lo32(double t) {
  return (int)(unsigned)(__builtin_bit_cast(unsigned long long, t) & 0xffffffffull);
}

// Digit scale of a line: v = rint(x SIGMA 2^(30 - E)), the exact product rounded once; a product
// of two lines comes back with UNSCALE = 255^3 / SIGMA^2 (rounded to FP64 once, the same constant
// as the format's model).
constexpr double SIGMA = 1.96875;
constexpr double UNSCALE = 16581375.0 * 1024.0 / 3969.0;

// Balanced base-255 digits of x under exponent E (|x| < 2^E), d[0] the most significant. With the
// bias 255 * 8290000 (about 2^31) the running value is a 32-bit unsigned number (|v| < 2^31), and
// floor(u / 255) is the high half of u * 0x80808081 shifted by 7 (exact for every 32-bit u).
__device__ __forceinline__ void digits_of(double x, int E, int (&d)[4]) {
  int v = lo32(__builtin_fma(x, __builtin_ldexp(SIGMA, 30 - E), RND));
#pragma unroll
  for (int k = 3; k >= 1; k--) {
    const unsigned u = (unsigned)v + 127u + 2113950000u;
    const unsigned q = __umulhi(u, 0x80808081u) >> 7;
    d[k] = (int)(u - 255u * q) - 127;
    v = (int)q - 8290000;
  }
  d[0] = v;
}
// The low bytes of a, b, c, d as one word (three byte permutes)
__device__ __forceinline__ unsigned pack4p(int a, int b, int c, int d) {
  const unsigned lo = __builtin_amdgcn_perm((unsigned)b, (unsigned)a, 0x0c0c0400u);
  const unsigned hi = __builtin_amdgcn_perm((unsigned)d, (unsigned)c, 0x0c0c0400u);
  return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// v = rint(x s) as an integer, for s = 2^(30 - E) and |x| < 2^E (one fma)
__device__ __forceinline__ int scaled_int(double x, double s) { return lo32(__builtin_fma(x, s, RND)); }
// The digits of digits_of by FP64 arithmetic in 8 operations (the GEMM epilogue's form), for
// s = SIGMA 2^(30 - E) and |x| < 2^E: with v = rint(x s) and Q_k = round(v / 255^k),
//   d = (Q_3, Q_2 - 255 Q_3, Q_1 - 255 Q_2, v - 255 Q_1).
// v / 255^(k+1) = Q_k / 255 + (v / 255^k - Q_k) / 255; the second term is below 1/510 in magnitude and
// Q_k / 255 is at least 1/510 away from a half-integer (255 is odd), so Q_(k+1) = round(Q_k / 255):
// each difference is a balanced digit, and these are the unique digits. v / 255^k itself is at least
// 1/(2 255^k) from a half-integer, far beyond the error of the product with the rounded 1/255^k
// (below 2^-44); |Q_1| <= 2^31 / 255 < 2^23, so the multiples of 255 are 24-bit products.
__device__ __forceinline__ void digits_fast(double x, double s, int (&d)[4]) {
  const double t0 = __builtin_fma(x, s, RND);
  const double v = t0 - RND;
  const int q1 = lo32(__builtin_fma(v, 1.0 / 255.0, RND));
  const int q2 = lo32(__builtin_fma(v, 1.0 / 65025.0, RND));
  const int q3 = lo32(__builtin_fma(v, 1.0 / 16581375.0, RND));
  d[0] = q3;
  d[1] = q2 + __mul24(q3, -255);
  d[2] = q1 + __mul24(q2, -255);
  d[3] = lo32(t0) + __mul24(q1, -255);
}

// Closing-factor form of a P x P block: each value as one int32, v = rint(x 2^(30-E)) under the same
// line exponent E as its digits (the same block size, 8 bytes per complex element). A 32 x 32
// chunk (rows 32 R.., columns 32 C..) is 8 KB at ((C P / 32 + R) * 8 KB); inside it element (r, j),
// with j = 8 g + 4 h + 2 p + e (g < 4, h, p, e < 2), sits at (2 g + p) KB + (32 h + r) * 16 + 8 e
// bytes, so that load k of lane l, 16 bytes at k KB + 16 l, holds columns 8 (k / 2) + 4 (l / 32) +
// 2 (k % 2), + 1 of row l % 32: the entries q = 2 k, 2 k + 1 of the lane's accumulator tile (row
// l % 32, columns 8 (q / 4) + 4 (l / 32) + q % 4).
__device__ __forceinline__ long long cf_off(int r, int j, int P) {
  const int jj = j & 31, rr = r & 31;
  return ((long long)(j >> 5) * (P >> 5) + (r >> 5)) * 8192 + (2 * (jj >> 3) + ((jj >> 1) & 1)) * 1024 +
         (32 * ((jj >> 2) & 1) + rr) * 16 + 8 * (jj & 1);
}

// Stored block of view block (x, y); its stored block row is the result / 4.
__device__ __forceinline__ int stored_block(const LcView& v, int x, int y) {
  if (v.kind == LC_KRON) return 0;
  const int ra = v.perm_r[x], cb = v.perm_c[y];
  return (v.flags & LC_TS) ? (cb * 4 + ra) : (ra * 4 + cb);
}
__device__ __forceinline__ void block_phase(const LcView& v, int x, int y, double& pr, double& pi) {
  if (v.kind == LC_KRON) {
    pr = 1.0;
    pi = 0.0;
    return;
  }
  pr = v.ph_r[x][0] * v.ph_c[y][0] - v.ph_r[x][1] * v.ph_c[y][1];
  pi = v.ph_r[x][0] * v.ph_c[y][1] + v.ph_r[x][1] * v.ph_c[y][0];
}

// x of lane l plus x of lane l ^ 32, then of l ^ 16: every lane holds the sum over its quad
// {l, l ^ 16, l ^ 32, l ^ 48}. A swap of a register with itself leaves the pair (x_l, x_partner) in
// the two results (their order depends on the lane half; the sum does not).
__device__ __forceinline__ double quad_sum(double x) {
  auto both = [](double v, auto swp) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    const auto l2 = swp(lo);
    const auto h2 = swp(hi);
    const double a = __builtin_bit_cast(double, ((unsigned long long)h2[0] << 32) | l2[0]);
    const double b = __builtin_bit_cast(double, ((unsigned long long)h2[1] << 32) | l2[1]);
    return a + b;
  };
  x = both(x, [](unsigned w) { return __builtin_amdgcn_permlane32_swap(w, w, false, false); });
  return both(x, [](unsigned w) { return __builtin_amdgcn_permlane16_swap(w, w, false, false); });
}

// A 64-bit value held in lane l of every wave, as a uniform (scalar) value
__device__ __forceinline__ long long readlane64(long long v, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

constexpr int TSTR = 32 * 16 + 16;  // LDS row stride of an FP64 epilogue tile (bytes)
constexpr int MAXA = 4;             // A chunks per step at most (dense x elemental: one per block column)

template <int NW>
constexpr int lds_bytes() {
  return (2 * MAXA * 16 * 1024 > NW * 32 * TSTR) ? 2 * MAXA * 16 * 1024 : NW * 32 * TSTR;
}

// Persistent schedule. The launch holds a few workgroups per XCD (as many as fit at once); XCD x
// owns the contiguous range of logical workgroups [x per, (x + 1) per), per = ceil(total / 8), and
// its j-th workgroup walks j, j + nj, j + 2 nj, ... of that range (nj workgroups per XCD). A
// logical workgroup is one strip of one item. As nj is a multiple of the workgroups per item, a
// workgroup keeps its strip from item to item: the per-strip tables (view blocks, phases, trace
// blocks) are built when the product group changes, and the first operand loads of the next item
// are issued before the epilogue of the current one.
// K256: every K sum of the launch has at most 256 terms (P <= 64), which lets the epilogue combine
// the two leading accumulators in 24-bit integer arithmetic. RUN: the launch's groups are runs of
// items sharing A (LcOzGroup::arun); without it the run logic is compiled out. XTR: the extended
// trace path, up to LC_OZ_MAX_TRACES traces per product and elemental closing factors (X (x) S);
// without it at most LC_MAX_FUSED_TRACES dense ones, with smaller tables and per-lane partials.
// CHAIN: the launch has chained groups (LcOzGroup::pre; not with RUN or XTR).
template <int NW, bool K256, bool RUN, bool XTR, bool CHAIN>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
lc_gemm_oz_kernel(const LcOzGroup* __restrict__ groups, int ng, const LcOzGroup* __restrict__ twins, int nv, int P,
                  long long total) {
  // stage buffers during the K loop, [2][MAXA] chunks of 16 planes x 1 KB (the 8 digit planes of
  // an A chunk and their negations), [nstep <= 8][1] for a run of dense products, or [4][2] for a
  // dense x elemental product at P = 128; the per-wave FP64 epilogue tiles overlap them
  __shared__ __attribute__((aligned(16))) unsigned char lds[lds_bytes<NW>()];
  __shared__ int rowe[NW][32];
  // partial sums of the fused traces (re, im), reduced by one wave per value: one per lane, or with
  // XTR one per lane quad {l, l + 16, l + 32, l + 48} (summed with cross-lane moves first)
  constexpr int MAXT = XTR ? LC_OZ_MAX_TRACES : LC_MAX_FUSED_TRACES;
  constexpr int TSL = XTR ? 16 : 64;  // partial sums per wave
  __shared__ d2v tsum[MAXT][NW * TSL];
  // per-wave tables of the current product group and strip (filled at a group change, read with
  // uniform LDS loads; in registers they would stay live through the K loop): A offsets, and per
  // sub-product (0: the group's product, 1: its twin) the B offsets, plane selectors and traces
  struct WaveTab {
    long long boff[4];
    double tfr[MAXT], tfi[MAXT];
    int ty[MAXT], te[MAXT];
    int flg[4], tfl[MAXT];
  };
  __shared__ WaveTab wtab[2][NW];
  __shared__ long long aofft[NW][4];
  // a chained group's first product (elemental x dense, LcOzGroup::pre): this wave's B block
  // offset and A plane selectors
  struct PreTab {
    long long boff;
    int flg, pad;
  };
  __shared__ PreTab ptab[NW];

  const int xcd = blockIdx.x & 7, nj = gridDim.x >> 3;
  const long long per = (total + 7) >> 3;
  const long long wend = (xcd + 1) * per < total ? (xcd + 1) * per : total;
  long long wg = xcd * per + (blockIdx.x >> 3);
  if (wg >= wend) return;

  // the wave index is uniform: readfirstlane lets the compiler keep everything derived from it in
  // scalar registers
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const long long P2 = (long long)P * P;
  const long long nv2 = (long long)nv * nv;
  const int ksteps = P >> 5;
  const int nstrip = (4 * P) >> 5;

  // A: one wave-wide 16-byte load per (chunk, plane), distributed over the waves, held in registers
  // for one step and then written to LDS. Lane l writes LDS bytes 16 l .. 16 l + 15 of the plane; it
  // fetches line 8 (l / 16) + (l % 16) / 2, K half (l % 2) ^ ((l / 16) & 1), so that each 16-lane
  // quarter reads 256 contiguous bytes and a fragment read (lane f: line f % 32, K half f / 32)
  // touches 64 distinct banks per quarter. (Register staging rather than direct-to-LDS loads: the
  // compiler cannot tell an LDS read from an LDS load in flight and would wait for all memory.)
  // Lane offsets are recomputed at the top of every item from an opaque copy of the lane index, so
  // that the compiler does not hoist their many derived addresses out of the item loop (they would
  // stay live across it and spill).
  int ln = lane, dma_off = 0, frag_off = 0, mem_off = 0, w_ = 0, h_ = 0;
  auto lane_consts = [&]() {
    ln = lane;
    asm volatile("" : "+v"(ln));
    const int lq = ln >> 4;
    dma_off = (8 * lq + ((ln & 15) >> 1)) * 32 + 16 * ((ln & 1) ^ (lq & 1));
    w_ = ln & 31;
    h_ = ln >> 5;
    frag_off = 16 * (16 * (w_ >> 3) + 2 * (w_ & 7) + (h_ ^ ((w_ >> 3) & 1)));
    mem_off = 32 * w_ + 16 * h_;  // the same fragment in memory order (B loads, digit stores)
  };

  // ---- per-(group, strip) state
  int gi = find_group(groups, ng, wg);
  int kind = 0, nseg = 1, wpi_log = 0, rr = 0, a = 0, i0 = 0, c = 0, j0 = 0, nA = 1, sA = 0, nstep = 0;
  int twi = -1;                     // the group's twin in `twins`, or -1
  int cpi = -1;                     // a chained group's first product in `twins`, or -1
  bool conjB_p = false;             // (its B conjugated)
  int ea_blk_p = 0, eb_blk_p = 0;   // (its exponent offsets, as ea_blk and eb_blk_s)
  int ea_chain = 0;                 // the row exponent of the chained A (the first product's output)
  int arun = 1;                     // items per logical workgroup (a run sharing its staged A)
  int mi = 0;                       // item of the run in progress
  int ntr_s[2] = {0, 0};            // fused traces per sub-product
  bool ykron_s[2] = {false, false};  // and whether any of their closing factors is elemental
  bool conjB_s[2] = {false, false};
  int ea_blk = 0, eb_blk_s[2] = {0, 0};  // exponent offsets of this strip's A rows and wave's B columns
  // Per K block b (kind 0: b < 4; kinds 1 and 2: b = 0) and A chunk (kind 2: slot s < nA, else the
  // chunk of block b), in aofft[wave] and wtab[sub][wave] (written by lanes 0..3): stored-block
  // byte offsets and this wave's A plane selectors. The product of a block pair is ph a b with
  // ph = i^k the pair's phase, a = Ar + i sA Ai (sA = -1: conjugated A) and b = Br + i sB Bi; with
  // w = ph a and B's imaginary part taken as sB Bi,
  //   re += Br re(w) - (sB Bi) im(w),   im += Br im(w) + (sB Bi) re(w),
  // where re(w) and im(w) are each one of Ar, Ai, -Ar, -Ai: the stage buffer holds all four (plane
  // group 0..3), and the selectors (bits 0-1: re(w), bits 2-3: im(w)) pick them, so the K loop
  // moves the B fragments to the matrix cores untouched (a conjugated B: one negation per digit).
  // Per fused trace t (written by lane 8 + t): the
  // closing factor's block (c, a) (FP64: byte offset of the block from the view base; closing-factor
  // form: of the stored block and of its row exponents), its read flags (bit 0 LC_TV, bit 1 LC_CONJ,
  // bit 2 a zero block) and the trace's factor, alpha_t times the block phase (zero for a zero
  // block). Item offsets are added per item. A twin product shares the group's A, kind and items.
  auto rfl = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  auto rfl64 = [](long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
  };
  auto tables = [&]() {
    const LcOzGroup& g = groups[gi];
    kind = g.kind;
    nseg = g.nseg;
    twi = g.twin;
    cpi = CHAIN ? g.pre : -1;
    arun = RUN ? g.arun : 1;
    wpi_log = 31 - __builtin_clz((unsigned)(nstrip * nseg));
    rr = (int)((wg - g.wg_begin) & ((1LL << wpi_log) - 1));
    const int strip = rr / nseg, seg = rr - strip * nseg;
    a = (strip * 32) / P;
    i0 = (strip * 32) % P;
    const int wcol0 = seg * NW * 32;  // first D-column of the workgroup
    c = (wcol0 + wave * 32) / P;      // this wave's block column
    j0 = (wcol0 + wave * 32) % P;     // and first column inside it
    const int cfirst = wcol0 / P;
    nstep = (kind == 0 ? 4 : 1) * ksteps;
    nA = kind == 2 ? (NW * 32) / P : 1;  // A chunks per step (dense x elemental: one per block column)
    sA = kind == 2 ? (wave * 32) / P : 0;
    const bool conjA = (g.A.flags & LC_CONJ) != 0;
    // view blocks of A slot s, and of this wave's B, at spin index b of the product
    auto a_block = [&](int s, int b, int& x, int& y) {
      x = a;
      y = kind == 2 ? cfirst + s : (kind == 1 ? a : b);
    };
    auto b_block = [&](int b, int& x, int& y) {
      if (kind == 2) {
        x = 0;
        y = 0;
        return;
      }
      x = kind == 1 ? a : b;
      y = c;
    };
    {
      int x, y;
      a_block(sA, 0, x, y);
      ea_blk = (stored_block(g.A, x, y) >> 2) * g.A.ebs + i0;
    }
    if (lane < 4) {
      long long aoff_l = 0;
      const int idx = lane;
      int x, y;
      if (kind == 2) {
        x = a;
        y = cfirst + idx;
      } else {
        a_block(0, idx, x, y);
      }
      if (y < 4) aoff_l = 8 * stored_block(g.A, x, y) * P2;
      aofft[wave][idx] = aoff_l;
    }
    if (cpi >= 0) {
      // the first product: elemental A1 (block (a, a)) times dense B1 (block (a, c)), K = P
      const LcOzGroup& gp = twins[cpi];
      conjB_p = (gp.B.flags & LC_CONJ) != 0;
      ea_blk_p = (stored_block(gp.A, a, a) >> 2) * gp.A.ebs + i0;
      eb_blk_p = (stored_block(gp.B, a, c) >> 2) * gp.B.ebs + j0;
      if (lane == 0) {
        double p1r, p1i, p2r, p2i;
        block_phase(gp.A, a, a, p1r, p1i);
        block_phase(gp.B, a, c, p2r, p2i);
        const double phr = p1r * p2r - p1i * p2i, phi = p1r * p2i + p1i * p2r;
        const int k = phi == 0.0 ? (phr > 0.0 ? 0 : 2) : (phi > 0.0 ? 1 : 3);
        const int gai = (gp.A.flags & LC_CONJ) ? 3 : 1;
        const int sre = k == 0 ? 0 : k == 1 ? (gai ^ 2) : k == 2 ? 2 : gai;
        const int sim = k == 0 ? gai : k == 1 ? 0 : k == 2 ? (gai ^ 2) : 2;
        ptab[wave].boff = 8 * stored_block(gp.B, a, c) * P2;
        ptab[wave].flg = sre | (sim << 2);
      }
    }
    for (int sub = 0; sub < (twi >= 0 ? 2 : 1); sub++) {
      const LcOzGroup& gs = sub ? twins[twi] : g;
      const int ntr = gs.ntr;
      ntr_s[sub] = ntr;
      bool yk = false;
      for (int t = 0; t < ntr; t++) yk = yk || gs.Y[t].kind == LC_KRON;
      ykron_s[sub] = yk;
      conjB_s[sub] = (gs.B.flags & LC_CONJ) != 0;
      {
        int x, y;
        b_block(0, x, y);
        eb_blk_s[sub] = (stored_block(gs.B, x, y) >> 2) * gs.B.ebs + j0;
      }
      WaveTab& wt = wtab[sub][wave];
      if (lane < 4) {
        const int idx = lane;
        int x, y;
        b_block(idx, x, y);
        wt.boff[idx] = 8 * stored_block(gs.B, x, y) * P2;
        int ax, ay;
        a_block(sA, idx, ax, ay);
        double p1r, p1i, p2r, p2i;
        block_phase(g.A, ax, ay, p1r, p1i);
        block_phase(gs.B, x, y, p2r, p2i);
        const double phr = p1r * p2r - p1i * p2i, phi = p1r * p2i + p1i * p2r;
        const int k = phi == 0.0 ? (phr > 0.0 ? 0 : 2) : (phi > 0.0 ? 1 : 3);  // ph = i^k
        // plane groups: 0 Ar, 1 Ai, 2 -Ar, 3 -Ai; sA Ai is group 1 (sA = 1) or 3
        const int gai = conjA ? 3 : 1;
        int sre, sim;
        if (k == 0) {
          sre = 0;
          sim = gai;
        } else if (k == 1) {  // i a = -sA Ai + i Ar
          sre = gai ^ 2;
          sim = 0;
        } else if (k == 2) {  // -a
          sre = 2;
          sim = gai ^ 2;
        } else {              // -i a = sA Ai - i Ar
          sre = gai;
          sim = 2;
        }
        wt.flg[idx] = sre | (sim << 2);
      }
      if (lane >= 8 && lane < 8 + ntr) {
        long long ty_l = 0, te_l = 0;
        const int t = lane - 8;
        const LcView& Yv = gs.Y[t];
        double pr, pi;
        bool zero = false;
        if (Yv.ex && Yv.kind == LC_KRON) {
          // X (x) S: block (c, a) is the phase times the one stored block where S has its non-zero
          zero = Yv.perm_r[c] != Yv.perm_c[a];
          pr = Yv.ph_r[c][0] * Yv.ph_c[a][0] - Yv.ph_r[c][1] * Yv.ph_c[a][1];
          pi = Yv.ph_r[c][0] * Yv.ph_c[a][1] + Yv.ph_r[c][1] * Yv.ph_c[a][0];
        } else if (Yv.ex) {
          const int blk = stored_block(Yv, c, a);
          ty_l = 8 * blk * P2;
          te_l = 4 * ((long long)(blk >> 2) * Yv.ebs);
          block_phase(Yv, c, a, pr, pi);
        } else {
          const BlockRef<double> Y = view_block<double>(Yv, 0, c, a, nv2);
          ty_l = (long long)(reinterpret_cast<const char*>(Y.p) - static_cast<const char*>(Yv.base));
          pr = Y.sr;
          pi = Y.si;
          zero = Y.zero;
        }
        const double ar = gs.tr_alpha[t][0], ai = gs.tr_alpha[t][1];
        wt.ty[t] = (int)ty_l;
        wt.te[t] = (int)te_l;
        wt.tfr[t] = zero ? 0.0 : ar * pr - ai * pi;
        wt.tfi[t] = zero ? 0.0 : ar * pi + ai * pr;
        wt.tfl[t] = ((Yv.flags & LC_TV) ? 1 : 0) | ((Yv.flags & LC_CONJ) ? 2 : 0) | (zero ? 4 : 0);
      }
    }
  };

  // ---- per-item state: batch coordinates and the operand byte offsets of the item
  struct ItemPos {
    int item, co[3];  // (items of a group count below 2^31)
    long long offA8, offB8;
    long long offA8p, offB8p;  // (a chained group's first product)
  };
  auto vo = [](const long long* st, const int* co) {
    return st[0] * co[0] + st[1] * co[1] + st[2] * co[2];
  };
  auto item_pos = [&](long long w) {
    const LcOzGroup& g = groups[gi];
    ItemPos q;
    q.item = (int)((w - g.wg_begin) >> wpi_log) * arun;
    const unsigned it = (unsigned)q.item, n0 = (unsigned)g.n[0], n1 = (unsigned)g.n[1];
    const unsigned q0 = it / n0;
    q.co[0] = (int)(it - q0 * n0);
    q.co[1] = (int)(q0 % n1);
    q.co[2] = (int)(q0 / n1);
    q.offA8 = 8 * vo(g.A.s, q.co);
    q.offB8 = 8 * vo(g.B.s, q.co);
    q.offA8p = q.offB8p = 0;
    if (cpi >= 0) {
      const LcOzGroup& gp = twins[cpi];
      q.offA8p = 8 * vo(gp.A.s, q.co);
      q.offB8p = 8 * vo(gp.B.s, q.co);
    }
    return q;
  };

  constexpr int LPW = 8 * MAXA / NW;  // A loads per wave and step at most
  i32x4 ra[LPW];
  i32x4 br[8];
  auto loadA = [&](int st, long long offA8) {
    const LcOzGroup& g = groups[gi];
    const int b = st / ksteps, kk = st - b * ksteps;
    const int nload = 8 * nA;
#pragma unroll
    for (int i = 0; i < LPW; i++) {
      const int f = wave + i * NW;
      if (i == 0 || f < nload) {  // (f < NW <= 8 <= nload for i = 0: every wave loads at least once)
        const int s = f >> 3, q = f & 7;
        const char* src = static_cast<const char*>(g.A.base) + offA8 + rfl64(aofft[wave][kind == 2 ? s : b]) +
                          q * P2 + ((long long)kk * P + i0) * 32;
        ra[i] = gload<i32x4>(src + dma_off);
      }
    }
  };
  // A of a chained group's first product (elemental: one chunk per step, at block (a, a))
  auto loadA_p = [&](int st, long long offA8) {
    const LcOzGroup& gp = twins[cpi];
#pragma unroll
    for (int i = 0; i < LPW; i++) {
      const int f = wave + i * NW;
      if (i == 0 || f < 8) {  // (f < NW <= 8 for i = 0: an unconditional load, as in loadA)
        const char* src = static_cast<const char*>(gp.A.base) + offA8 + (f & 7) * P2 + ((long long)st * P + i0) * 32;
        ra[i] = gload<i32x4>(src + dma_off);
      }
    }
  };
  // chunks between stage buffers: one in a run; nA for a dense x elemental product at P = 128 (its
  // 4 K steps of nA = 2 chunks fill the 8 slots); MAXA otherwise (at P <= 64 a dense x elemental
  // product has nA = MAXA, so its two steps fill them too)
  auto sstride = [&]() { return RUN ? 1 : (!K256 && kind == 2) ? nA : MAXA; };
  // plane q of chunk s (q < 4: Ar digit q, else Ai digit q - 4) into plane group q / 4, and its
  // negation into group q / 4 + 2; nc chunks (the product's nA, or 1 for a chained first product)
  auto storeA = [&](int bf, int nc) {
    const int nload = 8 * nc;
#pragma unroll
    for (int i = 0; i < LPW; i++) {
      const int f = wave + i * NW;
      if (f < nload) {
        unsigned char* d = &lds[((bf * sstride() + (f >> 3)) * 16 + (f & 7)) * 1024 + 16 * ln];
        *reinterpret_cast<i32x4*>(d) = ra[i];
        i32x4 n = ra[i];
        neg16(n);
        *reinterpret_cast<i32x4*>(d + 8 * 1024) = n;
      }
    }
  };
  // LDS-only barrier: the LDS writes before it are visible after it; memory loads in flight stay so
  auto lds_barrier = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  };
  // B of sub-product sub (2: a chained group's first product): this wave's fragments of step st come
  // straight from memory (plane q at + q P^2)
  auto bsrc = [&](int st, long long offB8, int sub) {
    if (sub == 2)
      return static_cast<const char*>(twins[cpi].B.base) + offB8 + rfl64(ptab[wave].boff) +
             ((long long)st * P + j0) * 32 + mem_off;
    const LcOzGroup& gs = sub ? twins[twi] : groups[gi];
    const int b = st / ksteps, kk = st - b * ksteps;
    return static_cast<const char*>(gs.B.base) + offB8 + rfl64(wtab[sub][wave].boff[b]) +
           ((long long)kk * P + j0) * 32 + mem_off;
  };
  // B of step 0, issued in the order the reloads use (digit by digit, real then imaginary)
  auto loadB0 = [&](long long offB8, int sub) {
    const char* src = bsrc(0, offB8, sub);
#pragma unroll
    for (int d = 0; d < NORD; d++) {
      br[d] = gload<i32x4>(src + d * P2);
      br[4 + d] = gload<i32x4>(src + (4 + d) * P2);
    }
  };
  // the first loads of an item: A of step 0 and B of step 0 (of the first product when chained)
  auto prologue = [&](const ItemPos& q) {
    if (CHAIN && cpi >= 0) {
      loadA_p(0, q.offA8p);
      loadB0(q.offB8p, 2);
    } else {
      loadA(0, q.offA8);
      loadB0(q.offB8, 0);
    }
  };

  lane_consts();
  tables();
  ItemPos cur = item_pos(wg);
  prologue(cur);

  for (;;) {
    lane_consts();
    const LcOzGroup& g = groups[gi];
    // next item of this workgroup (decided in the item's last sub-product)
    long long nwg = 0;
    bool more_items = false, pre = false;
    int ngi = gi;
    ItemPos nxt;  // (set when pre)
    // sub-product 0 is the group's product; 1, its twin, reuses the A operand staged for 0 (the
    // twin needs every step of A in the stage buffers: a dense x elemental product); 2, a chained group's first
    // product, runs before 0 and leaves its output in the stage buffers as 0's A (digits and row
    // exponents as a digit output would have them). One code path per sub-product.
    auto run_sub = [&](auto SUB) {
      constexpr int sub = decltype(SUB)::value;
      constexpr bool FIRST = sub == 2;
      const LcOzGroup& gs = FIRST ? twins[cpi] : sub ? twins[twi] : g;
      const int ntr = FIRST ? 0 : ntr_s[sub & 1];
      const bool conjB = FIRST ? conjB_p : conjB_s[sub & 1];
      const long long offB8 = FIRST ? cur.offB8p : sub ? 8 * vo(gs.B.s, cur.co) : cur.offB8;
      const int nst = FIRST ? ksteps : nstep;  // (the first product: elemental x dense, K = P)
      i32x16 re[NORD], im[NORD];
#pragma unroll
      for (int o = 0; o < NORD; o++) {
        re[o] = i32x16{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        im[o] = re[o];
      }
      // One K step: A from stage buffer bf, B from br. B digit db's two planes are refreshed for the
      // next step right after their last use here, so every reload has a whole step to land and one
      // register set suffices (the compiler waits for each plane before its first use).
      auto compute = [&](int st, int bf, bool more) {
        const int fl = FIRST ? rfl(ptab[wave].flg) : rfl(wtab[sub & 1][wave].flg[st / ksteps]);
        const int ore = (fl & 3) * 4096, oim = ((fl >> 2) & 3) * 4096;  // plane groups of re(w), im(w)
        const int onim = oim ^ 8192;                                     // and of -im(w)
        const char* nsrc = more ? bsrc(st + 1, offB8, sub) : nullptr;
        const unsigned char* asrc = &lds[((bf * sstride() + (FIRST ? 0 : sA)) * 16) * 1024 + frag_off];
        auto rd = [&](int off) { return *reinterpret_cast<const i32x4*>(asrc + off); };
        // one digit pair: re += Br re(w) - (sB Bi) im(w), im += Br im(w) + (sB Bi) re(w); re is reused
        // two MFMAs later
        auto pair = [&](int o, const i32x4& xr, const i32x4& xi, const i32x4& w, const i32x4& wi, const i32x4& nw) {
          re[o] = mfma_i8(xr, w, re[o]);
          im[o] = mfma_i8(xr, wi, im[o]);
          re[o] = mfma_i8(xi, nw, re[o]);
          im[o] = mfma_i8(xi, w, im[o]);
        };
        // A digit 0 meets every B digit: its three fragments stay in registers for the step; the
        // fragments of the other pairs are read one pair ahead of their products
        const i32x4 z0 = rd(ore), z1 = rd(oim), z2 = rd(onim);
#pragma unroll
        for (int db = 0; db < NORD; db++) {
          // keeps the scheduler from hoisting digit db's use of B above the previous digits' products:
          // each plane then waits only for its own reload, a whole step after it was issued (LDS
          // reads may still move across)
          __builtin_amdgcn_sched_barrier(0x0100);
          const i32x4 xr = br[db];
          i32x4 xi = br[4 + db];
          if (conjB) neg16(xi);
          i32x4 c0, c1, c2;
          if (db + 1 < NORD) {
            c0 = rd(ore + 1024);
            c1 = rd(oim + 1024);
            c2 = rd(onim + 1024);
          }
          pair(db, xr, xi, z0, z1, z2);
#pragma unroll
          for (int da = 1; da + db < NORD; da++) {
            // (compile-time after unrolling) the fragments of the next pair, if any
            const bool nxt = da + 1 + db < NORD;
            i32x4 n0, n1, n2;
            if (nxt) {
              n0 = rd(ore + (da + 1) * 1024);
              n1 = rd(oim + (da + 1) * 1024);
              n2 = rd(onim + (da + 1) * 1024);
            }
            pair(da + db, xr, xi, c0, c1, c2);
            if (nxt) {
              c0 = n0;
              c1 = n1;
              c2 = n2;
            }
          }
          if (more) {
            br[db] = gload<i32x4>(nsrc + db * P2);
            br[4 + db] = gload<i32x4>(nsrc + (4 + db) * P2);
          }
        }
      };

      // The loop runs every step but the last with unconditional loads, so that the compiler's count
      // of loads in flight is exact on every path and each wait covers only what it needs. Stage
      // buffer of step st: st in a run (every step kept for the run's later items) and for a dense x
      // elemental product at P = 128 (4 steps of 2 chunks: every step kept for a twin), else st % 2.
      auto sbuf = [&](int st) { return (RUN || (!K256 && kind == 2)) ? st : (st & 1); };
      if (FIRST) {
        storeA(0, 1);
        lds_barrier();
        for (int st = 0; st + 1 < nst; st++) {
          loadA_p(st + 1, cur.offA8p);
          compute(st, st & 1, true);
          storeA((st + 1) & 1, 1);
          lds_barrier();
        }
      } else if (sub == 0 && (!RUN || mi == 0) && cpi < 0) {
        storeA(0, nA);
        lds_barrier();
        for (int st = 0; st + 1 < nstep; st++) {
          // A of step st + 1 is loaded during step st and written to the other stage buffer after it;
          // one barrier per step suffices (that buffer was last read in step st - 1)
          loadA(st + 1, cur.offA8);
          compute(st, sbuf(st), true);
          storeA(sbuf(st + 1), nA);
          lds_barrier();
        }
      } else {
        // the twin, a later item of a run, or a chained product: every step of A is in the stage
        // buffers
        for (int st = 0; st + 1 < nstep; st++) compute(st, sbuf(st), true);
      }
      // exponents of the epilogue, loaded before the last step: row i0 + lane % 32 of A; columns
      // j0 + 8 g + 4 (lane / 32) + r of B
      const int ea = FIRST ? gload<int>(gs.A.ex + vo(gs.A.es, cur.co) + ea_blk_p + w_)
                     : cpi >= 0 ? ea_chain
                                : gload<int>(g.A.ex + vo(g.A.es, cur.co) + ea_blk + w_);
      int eb[16];
      {
        const int* pe = gs.B.ex + vo(gs.B.es, cur.co) + (FIRST ? eb_blk_p : eb_blk_s[sub & 1]) + 4 * h_;
#pragma unroll
        for (int gg = 0; gg < 4; gg++) {
          const i32x4 v = gload<i32x4>(pe + 8 * gg);
#pragma unroll
          for (int r = 0; r < 4; r++) eb[4 * gg + r] = v[r];
        }
      }
      compute(nst - 1, FIRST ? ((nst - 1) & 1) : sbuf(nstep - 1), false);
      // a chained group: the product's first B fragments, while the first product's epilogue runs
      if constexpr (FIRST) loadB0(cur.offB8, 0);

      // ---- epilogue
      // S = ((acc0 255 + acc1) 255 + acc2) 255 + acc3, exact in FP64 (|acc_o| < 2^27, |S| < 2^51);
      // the value is UNSCALE 2^(E_A + E_B - 60) S, rounded once. For K <= 256 (P <= 64) the first
      // step is integer: |acc0| <= 2 K 127^2 < 2^23 (a 24-bit product) and |acc0 255 + acc1| < 2^31.
      auto comb = [](int a0, int a1, int a2, int a3) {
        const double h = K256 ? (double)(__mul24(a0, 255) + a1) : __builtin_fma((double)a0, 255.0, (double)a1);
        return __builtin_fma(__builtin_fma(h, 255.0, (double)a2), 255.0, (double)a3);
      };
      double vr[16], vi[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const double sc = __builtin_ldexp(UNSCALE, ea + eb[q] - 60);
        vr[q] = comb(re[0][q], re[1][q], re[2][q], re[3][q]) * sc;
        vi[q] = comb(im[0][q], im[1][q], im[2][q], im[3][q]) * sc;
      }
      if (!(gs.alpha[0] == 1.0 && gs.alpha[1] == 0.0)) {
        const double alr = gs.alpha[0], ali = gs.alpha[1];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const double xr = vr[q], xi = vi[q];
          vr[q] = alr * xr - ali * xi;
          vi[q] = alr * xi + ali * xr;
        }
      }
      // digit output (and a chained group's first product): exponent of the D-row = the largest
      // |value| over every wave of the strip (the whole row at P <= 64, half of it at P = 128),
      // exchanged through LDS
      const bool dig = FIRST || gs.Cd;
      if (dig) {
        double mx = 0.0;
#pragma unroll
        for (int q = 0; q < 16; q++) mx = __builtin_fmax(mx, __builtin_fmax(__builtin_fabs(vr[q]), __builtin_fabs(vi[q])));
        int em = exp_of(mx);
        const int eo = __shfl_xor(em, 32, 64);
        em = eo > em ? eo : em;
        if (lane < 32) rowe[wave][lane] = em;
      }
      // The FP64 tile in LDS serves only FP64 outputs and FP64 closing factors; traces in the
      // closing-factor form, closing-factor outputs and digit outputs work on the lane's registers.
      const bool use_tile = !FIRST && (gs.C || (ntr > 0 && !gs.Y[0].ex));
      // every wave is past the K loop (the stage buffers are free for the epilogue tiles) and the row
      // exponents are visible
      if (dig || use_tile) lds_barrier();
      int E = 0;
      if (dig) {
        E = INT_MIN;
#pragma unroll
        for (int w = 0; w < NW; w++) {
          const int x = rowe[w][w_];
          E = x > E ? x : E;
        }
        if constexpr (!K256) {
          // half a D-row (P = 128, the strip's columns 256 h ..): its exponent as it is (INT_MIN when
          // the half is zero) at [D-row][h], the digits under it; lc_oz_fix_kernel makes the row's
          if (wave == 0 && lane < 32)
            gstore(gs.Ce + vo(gs.ces, cur.co) + 2 * ((long long)a * P + i0 + lane) + ((c * P + j0) >> 8), E);
        }
        if (E == INT_MIN) E = 0;
        if (FIRST) ea_chain = E;
        else if (K256 && wave == 0 && lane < 32) gstore(gs.Ce + vo(gs.ces, cur.co) + (long long)a * P + i0 + lane, E);
      }
      const double sE = __builtin_ldexp(1.0, 30 - E);
      // this wave's tile in LDS as FP64 complex, row-major with a padded row stride; only this wave
      // reads it (LDS operations of one wave complete in order: no barrier)
      unsigned char* tile = &lds[wave * 32 * TSTR];
      const int hcol = 4 * h_;
      if (use_tile) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int col = 8 * (q >> 2) + hcol + (q & 3);
          *reinterpret_cast<d2v*>(tile + w_ * TSTR + col * 16) = d2v{vr[q], vi[q]};
        }
      }
      if (!FIRST && gs.C) {
        // whole rows: lanes 0..31 one row, 32..63 the next
        Cplx<double>* C = reinterpret_cast<Cplx<double>*>(gs.C) + vo(gs.cs, cur.co) + (long long)(a * 4 + c) * nv2;
#pragma unroll 4
        for (int m = 0; m < 16; m++) {
          const int lr = 2 * m + h_, lc = w_;
          const int grow = i0 + lr, gcol = j0 + lc;
          if (grow < nv && gcol < nv) {
            const d2v v = *reinterpret_cast<const d2v*>(tile + lr * TSTR + lc * 16);
            gstore_c(C + (long long)grow * nv + gcol, Cplx<double>{v[0], v[1]});
          }
        }
      }

      if constexpr (!FIRST) if (ntr > 0) {
        const int w = w_, h = h_;
        // the trace value of this lane into the partial sums: summed over its lane quad first, with
        // the lane swaps of v_permlane32_swap (lanes l, l ^ 32) and v_permlane16_swap (l, l ^ 16)
        auto put_tsum = [&](int t, double xr, double xi) {
          if constexpr (XTR) {
            xr = quad_sum(xr);
            xi = quad_sum(xi);
            if (lane < 16) tsum[t][wave * 16 + lane] = d2v{xr, xi};
          } else {
            tsum[t][tid] = d2v{xr, xi};
          }
        };
        // FP64 closing factors (one trace at a time: 64 registers of loads each): trace t reads block
        // (c, a) of f(Y_t), element (col, row): with LC_TV the stored element (row, col), read along
        // stored rows; otherwise the stored (col, row). Lane l reads stored row r0 + 2 m + l / 32,
        // column c0 + l % 32.
        const bool full = i0 + 32 <= nv && j0 + 32 <= nv;
        auto trace_fp64 = [&](int t) {
          Cplx<double> yv[16];
          const Cplx<double>* p = reinterpret_cast<const Cplx<double>*>(
              static_cast<const char*>(gs.Y[t].base) + 16 * vo(gs.Y[t].s, cur.co) + rfl(wtab[sub][wave].ty[t]));
          const int fl = rfl(wtab[sub][wave].tfl[t]);
          const bool tv = (fl & 1) != 0;
          const double cj = (fl & 2) ? -1.0 : 1.0;
          const int r0 = tv ? i0 : j0, c0 = tv ? j0 : i0;
          const Cplx<double>* q = p + (long long)(r0 + h) * nv + c0 + w;
          const long long rs = 2LL * nv;
          if (full) {
#pragma unroll
            for (int m = 0; m < 16; m++) yv[m] = gload_c(q + m * rs);
          } else {
#pragma unroll
            for (int m = 0; m < 16; m++) {
              const bool in = c0 + w < nv && r0 + 2 * m + h < nv;
              yv[m] = gload_c(in ? q + m * rs : p);
              if (!in) yv[m] = Cplx<double>{0.0, 0.0};
            }
          }
          double sr = 0, si = 0;
#pragma unroll
          for (int m = 0; m < 16; m++) {
            const int u = 2 * m + h;
            const int lr = tv ? u : w, lc = tv ? w : u;  // tile row / column
            const d2v x = *reinterpret_cast<const d2v*>(tile + lr * TSTR + lc * 16);
            const double yr = yv[m].x, yi = cj * yv[m].y;
            sr += x[0] * yr - x[1] * yi;
            si += x[0] * yi + x[1] * yr;
          }
          const double fr = __builtin_bit_cast(double, rfl64(__builtin_bit_cast(long long, wtab[sub][wave].tfr[t])));
          const double fi = __builtin_bit_cast(double, rfl64(__builtin_bit_cast(long long, wtab[sub][wave].tfi[t])));
          put_tsum(t, fr * sr - fi * si, fr * si + fi * sr);
        };
        // Closing factors in the closing-factor form (every trace of the product, LC_TV): block (c, a)
        // of f(Y) element (col, row) is the stored (row, col), so the wave's part is the chunk of its
        // C tile, laid out like the accumulators: load k of lane l holds the values of its own tile
        // entries q = 2 k and 2 k + 1 (row l % 32, columns 8 (k / 2) + 4 (l / 32) + 2 (k % 2), + 1),
        // under that row's exponent, so each lane sums FP64 products of its registers with exact
        // integers and scales once. Traces T0 .. T0 + N - 1, every load of them issued before the
        // first use (one memory round trip); with SKIP (elemental closing factors) a zero block of
        // the wave is skipped.
        auto traces_cf = [&](auto T0c, auto NT, auto SKIPc) {
          constexpr int T0 = decltype(T0c)::value, N = decltype(NT)::value;
          constexpr bool SKIP = decltype(SKIPc)::value;
          i32x4 yd[N][8];
          int ey[N], fl[N];
#pragma unroll
          for (int t = 0; t < N; t++) {
            fl[t] = rfl(wtab[sub][wave].tfl[T0 + t]);
            if (SKIP && (fl[t] & 4)) continue;
            const LcView& Yv = gs.Y[T0 + t];
            const unsigned char* yb = static_cast<const unsigned char*>(Yv.base) + 8 * vo(Yv.s, cur.co) +
                                      rfl(wtab[sub][wave].ty[T0 + t]) +
                                      ((long long)(j0 >> 5) * (P >> 5) + (i0 >> 5)) * 8192 + 16 * ln;
#pragma unroll
            for (int k = 0; k < 8; k++) yd[t][k] = gload<i32x4>(yb + 1024 * k);
            const char* eyb = reinterpret_cast<const char*>(Yv.ex + vo(Yv.es, cur.co)) + rfl(wtab[sub][wave].te[T0 + t]);
            ey[t] = gload<int>(reinterpret_cast<const int*>(eyb) + i0 + w);
          }
#pragma unroll
          for (int t = 0; t < N; t++) {
            double sr = 0.0, si = 0.0;
            if (!SKIP || !(fl[t] & 4)) {
              const double cj = (fl[t] & 2) ? -1.0 : 1.0;
              double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma unroll
              for (int k = 0; k < 8; k++) {
#pragma unroll
                for (int p2 = 0; p2 < 2; p2++) {
                  const double xr = vr[2 * k + p2], xi = vi[2 * k + p2];
                  const double yr = (double)yd[t][k][2 * p2], yi = (double)yd[t][k][2 * p2 + 1];
                  s1 += xr * yr;
                  s2 += xi * yi;
                  s3 += xr * yi;
                  s4 += xi * yr;
                }
              }
              // x (yr + i cj yi), scaled by the row's 2^(E - 30)
              const double ur = __builtin_ldexp(s1 - cj * s2, ey[t] - 30);
              const double ui = __builtin_ldexp(cj * s3 + s4, ey[t] - 30);
              const double fr = __builtin_bit_cast(double, rfl64(__builtin_bit_cast(long long, wtab[sub][wave].tfr[T0 + t])));
              const double fi = __builtin_bit_cast(double, rfl64(__builtin_bit_cast(long long, wtab[sub][wave].tfi[T0 + t])));
              sr = fr * ur - fi * ui;
              si = fr * ui + fi * ur;
            }
            put_tsum(T0 + t, sr, si);
          }
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        using NO = std::false_type;
        using YES = std::true_type;
        if (gs.Y[0].ex && (!XTR || !ykron_s[sub])) {
          if (ntr == 1) traces_cf(I0{}, I1{}, NO{});
          else if (ntr == 2) traces_cf(I0{}, I2{}, NO{});
          else traces_cf(I0{}, I3{}, NO{});
          // more than three traces: the product's own (a twin has at most three)
          if constexpr (XTR && sub == 0) {
            if (ntr == 4) traces_cf(I3{}, I1{}, NO{});
            else if (ntr == 5) traces_cf(I3{}, I2{}, NO{});
            else if (ntr == 6) traces_cf(I3{}, I3{}, NO{});
          }
        } else if (XTR && gs.Y[0].ex) {
          // elemental closing factors (X (x) S): most blocks of a wave are zero
          if (ntr == 1) traces_cf(I0{}, I1{}, YES{});
          else if (ntr == 2) traces_cf(I0{}, I2{}, YES{});
          else traces_cf(I0{}, I3{}, YES{});
          if constexpr (sub == 0) {
            if (ntr == 4) traces_cf(I3{}, I1{}, YES{});
            else if (ntr == 5) traces_cf(I3{}, I2{}, YES{});
            else if (ntr == 6) traces_cf(I3{}, I3{}, YES{});
          }
        } else {
          for (int t = 0; t < ntr; t++) trace_fp64(t);
        }
        lds_barrier();
        // wave v sums component k = v % 2 of trace t = v / 2 over the workgroup's TSL NW partial sums
        for (int v = wave; v < 2 * ntr; v += NW) {
          const int t = v >> 1, k = v & 1;
          const double* ts = reinterpret_cast<const double*>(&tsum[t][0]) + k;
          double sm = 0;
#pragma unroll
          for (int j = 0; j < NW * TSL / 64; j++) sm += ts[2 * (lane + 64 * j)];
          sm = wave_sum(sm);
          if (lane == 0) gstore(gs.part[t] + 2 * (((long long)cur.item << wpi_log) + rr) + k, sm);
        }
      }

      const bool last_sub = !FIRST && (sub == 1 || twi < 0);  // (uniform)
      if (FIRST) {
        // (the product's first B fragments went out after the last K step)
      } else if (RUN && last_sub && mi + 1 < arun) {
        // the next item of the run: its first B fragments (A is on the chip)
        const LcOzGroup& g0 = groups[gi];
        loadB0(cur.offB8 + 8 * g0.B.s[0], 0);
      } else if (last_sub) {
        // the next item of this workgroup: with the same tables (same group and strip), its first
        // loads go out now and land while this item's output is written
        nwg = wg + nj;
        more_items = nwg < wend;
        if (more_items)
          while (ngi + 1 < ng && groups[ngi + 1].wg_begin <= nwg) ngi++;
        pre = more_items && ngi == gi && (nj & ((1 << wpi_log) - 1)) == 0;
        if (pre) {
          nxt = item_pos(nwg);
          prologue(nxt);
        }
      } else {
        // the twin's first B fragments
        loadB0(8 * vo(twins[twi].B.s, cur.co), 1);
      }

      if (!FIRST && gs.Cd && gs.cf) {
        // closing-factor form: load k of the chunk holds the lane's own entries q = 2 k, 2 k + 1 (see
        // traces_cf); one coalesced kilobyte per k
        unsigned char* base = static_cast<unsigned char*>(gs.Cd) + 8 * (vo(gs.cds, cur.co) + (long long)(a * 4 + c) * P2) +
                              ((long long)(j0 >> 5) * (P >> 5) + (i0 >> 5)) * 8192 + 16 * ln;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const i32x4 v = {scaled_int(vr[2 * k], sE), scaled_int(vi[2 * k], sE), scaled_int(vr[2 * k + 1], sE),
                           scaled_int(vi[2 * k + 1], sE)};
          gstore(base + 1024 * k, v);
        }
      } else if (dig) {
        // digits of the lane's 16 values: word W[d][gg] packs digit plane d (re digits 0..3, then im)
        // of columns 8 gg + 4 (l / 32) .. + 3 of row l % 32. A stored plane row of a lane is 16
        // adjacent columns (16 (l / 32) ..): lane l keeps W[.][0..1] (l < 32) or W[.][2..3] and takes
        // the other two from lane l ^ 32, one v_permlane32_swap per pair (gg, gg + 2): the swap of
        // the upper lanes of x with the lower lanes of y leaves x = (W_gg of lanes 0..31, W_gg+2 of
        // lanes 0..31) and y = (W_gg of lanes 32..63, W_gg+2 of lanes 32..63), which are the words
        // 2 gg' and 2 gg' + 1 of both halves.
        unsigned wd[8][4];
        const double sD = sE * SIGMA;  // (exact)
#pragma unroll
        for (int gg = 0; gg < 4; gg++) {
          int dr[4][4], di[4][4];  // [column r][digit]
#pragma unroll
          for (int r = 0; r < 4; r++) {
            digits_fast(vr[4 * gg + r], sD, dr[r]);
            digits_fast(vi[4 * gg + r], sD, di[r]);
          }
#pragma unroll
          for (int d = 0; d < 4; d++) {
            wd[d][gg] = pack4p(dr[0][d], dr[1][d], dr[2][d], dr[3][d]);
            wd[4 + d][gg] = pack4p(di[0][d], di[1][d], di[2][d], di[3][d]);
          }
        }
        if constexpr (FIRST) {
          // as the chained product's A: chunk (K half j0 / 32, block column c) of the stage buffers,
          // whose lines are this strip's rows; the lane's 16 bytes of plane q (row l % 32, columns
          // 16 (l / 32) ..) go where storeA would put them (the fragment offset), with their negation
          unsigned char* base = &lds[(((j0 >> 5) * MAXA + c) * 16) * 1024 + frag_off];
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const auto s0 = __builtin_amdgcn_permlane32_swap(wd[q][0], wd[q][2], false, false);
            const auto s1 = __builtin_amdgcn_permlane32_swap(wd[q][1], wd[q][3], false, false);
            i32x4 v = {(int)s0[0], (int)s0[1], (int)s1[0], (int)s1[1]};
            *reinterpret_cast<i32x4*>(base + q * 1024) = v;
            neg16(v);
            *reinterpret_cast<i32x4*>(base + (8 + q) * 1024) = v;
          }
          // every chunk written before the chained product reads them
          lds_barrier();
        } else {
          unsigned char* base = static_cast<unsigned char*>(gs.Cd) + 8 * (vo(gs.cds, cur.co) + (long long)(a * 4 + c) * P2) +
                                ((long long)(j0 >> 5) * P + i0) * 32;
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const auto s0 = __builtin_amdgcn_permlane32_swap(wd[q][0], wd[q][2], false, false);
            const auto s1 = __builtin_amdgcn_permlane32_swap(wd[q][1], wd[q][3], false, false);
            const i32x4 v = {(int)s0[0], (int)s0[1], (int)s1[0], (int)s1[1]};
            gstore(base + q * P2 + mem_off, v);
          }
        }
      }

    };
    if constexpr (CHAIN) {
      // a chained group: its first product, whose output stays in the stage buffers
      if (cpi >= 0) run_sub(std::integral_constant<int, 2>{});
    }
    run_sub(std::integral_constant<int, 0>{});
    if (twi >= 0) {
      // the twin writes the trace sums that slower waves may still be reading
      lds_barrier();
      run_sub(std::integral_constant<int, 1>{});
    }
    if (RUN && mi + 1 < arun) {
      // the next item of the run (A staged once, with the first): it writes the row exponents and
      // trace sums that slower waves may still be reading
      lds_barrier();
      mi++;
      cur.item++;
      cur.co[0]++;
      cur.offB8 += 8 * g.B.s[0];
      continue;
    }
    mi = 0;

    if (!more_items) break;
    // every wave is done with its tile and the trace sums before the stage buffers are reused
    lds_barrier();
    wg = nwg;
    if (pre) {
      cur = nxt;
    } else {
      gi = ngi;
      tables();
      cur = item_pos(wg);
      prologue(cur);
    }
  }
}

// ------------------------------------------------------------------------------------------
// Conversion FP64 -> digits (inputs).
//
// Exponents: per matrix, one exponent per row index (over every block and column) into ex_r and one
// per column index (over every block and row) into ex_t, P ints each. nblk = 16 (D-space) or 1
// (elemental); matrices of nblk blocks of nv x nv complex doubles, `stride` complex elements apart.
// One workgroup per matrix.
__global__ void __launch_bounds__(256) lc_oz_exps_kernel(const double2* __restrict__ x, long long stride, int nv,
                                                         int P, int nblk, int* __restrict__ ex_r,
                                                         int* __restrict__ ex_t) {
  __shared__ int er[128], et[128];
  const long long m = blockIdx.x;
  for (int i = threadIdx.x; i < P; i += 256) {
    er[i] = INT_MIN;
    et[i] = INT_MIN;
  }
  __syncthreads();
  const double2* p = x + m * stride;
  const long long nv2 = (long long)nv * nv;
  const long long n = nblk * nv2;
  for (long long e = threadIdx.x; e < n; e += 256) {
    const int r = (int)((e / nv) % nv), cc = (int)(e % nv);
    const double2 v = p[e];
    const int e1 = exp_of(v.x), e2 = exp_of(v.y);
    const int ev = e1 > e2 ? e1 : e2;
    if (ev == INT_MIN) continue;
    atomicMax(&er[r], ev);
    atomicMax(&et[cc], ev);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += 256) {
    ex_r[m * P + i] = er[i] == INT_MIN ? 0 : er[i];
    ex_t[m * P + i] = et[i] == INT_MIN ? 0 : et[i];
  }
}

// Digits. One workgroup per 32 x 32 tile of a P x P block (zero outside nv x nv): the row-line
// planes (line = row, K = column) and the column-line planes (line = column, K = row), one
// contiguous kilobyte per plane and tile; the bytes are gathered in LDS first. Optionally the
// closing-factor form under the row exponents (ycf), and that of the transposed block under the
// column exponents (ycft: element (j, i) of X^T at its place).
__global__ void __launch_bounds__(256) lc_oz_digits_kernel(const double2* __restrict__ x, long long stride, int nv,
                                                           int P, int nblk, const int* __restrict__ ex_r,
                                                           const int* __restrict__ ex_t, unsigned char* __restrict__ yr,
                                                           unsigned char* __restrict__ yt, int2* __restrict__ ycf,
                                                           int2* __restrict__ ycft) {
  __shared__ __attribute__((aligned(16))) unsigned char tr_[8][32][32], tt_[8][32][32];
  const int tpb = P >> 5;
  const long long t = blockIdx.x;
  const long long bi = t / (tpb * tpb);  // global block index: matrix * nblk + block
  const long long m = bi / nblk;
  const int blk = (int)(bi - m * nblk);
  const int tr = (int)((t / tpb) % tpb), tc = (int)(t % tpb);
  const double2* p = x + m * stride + (long long)blk * nv * nv;
  const long long P2 = (long long)P * P;
#pragma unroll
  for (int pass = 0; pass < 4; pass++) {
    const int r = (threadIdx.x >> 5) + 8 * pass, cc = threadIdx.x & 31;
    const int row = tr * 32 + r, col = tc * 32 + cc;
    double2 v = make_double2(0.0, 0.0);
    if (row < nv && col < nv) v = p[(long long)row * nv + col];
    int dr[4], di[4];
    const int er = ex_r[m * P + row], et = ex_t[m * P + col];
    if (ycf)
      ycf[(bi * 8 * P2 + cf_off(row, col, P)) / 8] =
          make_int2((int)__builtin_rint(__builtin_ldexp(v.x, 30 - er)), (int)__builtin_rint(__builtin_ldexp(v.y, 30 - er)));
    if (ycft)
      ycft[(bi * 8 * P2 + cf_off(col, row, P)) / 8] =
          make_int2((int)__builtin_rint(__builtin_ldexp(v.x, 30 - et)), (int)__builtin_rint(__builtin_ldexp(v.y, 30 - et)));
    digits_of(v.x, er, dr);
    digits_of(v.y, er, di);
#pragma unroll
    for (int d = 0; d < 4; d++) {
      tr_[d][r][cc] = (unsigned char)dr[d];
      tr_[4 + d][r][cc] = (unsigned char)di[d];
    }
    digits_of(v.x, et, dr);
    digits_of(v.y, et, di);
#pragma unroll
    for (int d = 0; d < 4; d++) {
      tt_[d][cc][r] = (unsigned char)dr[d];
      tt_[4 + d][cc][r] = (unsigned char)di[d];
    }
  }
  __syncthreads();
  // 8 planes x 1 KB each: 512 16-byte pieces, two per thread
  unsigned char* ob = yr + bi * 8 * P2;
  unsigned char* obt = yt + bi * 8 * P2;
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const int k = threadIdx.x + 256 * pass;
    const int q = k >> 6, piece = k & 63;
    const int ln = piece >> 1, half = piece & 1;
    const uint4 v = *reinterpret_cast<const uint4*>(&tr_[q][ln][16 * half]);
    *reinterpret_cast<uint4*>(ob + q * P2 + ((long long)tc * P + tr * 32 + ln) * 32 + 16 * half) = v;
    const uint4 w = *reinterpret_cast<const uint4*>(&tt_[q][ln][16 * half]);
    *reinterpret_cast<uint4*>(obt + q * P2 + ((long long)tr * P + tc * 32 + ln) * 32 + 16 * half) = w;
  }
}

// ------------------------------------------------------------------------------------------
// Row exponents of a P = 128 digit output. A product's workgroup strip there covers half a D-row
// (block columns 2 h, 2 h + 1), so the GEMM writes each half's digits (or closing-factor values)
// under the half's own exponent, eh[row][h] (INT_MIN for an all-zero half). This kernel makes the
// D-row's: E = max(eh[row][0], eh[row][1]) into ex[row] (0 for an all-zero row, as the input
// conversion gives it), and the half with the smaller exponent, k = E - eh[row][h] > 0, is
// rescaled in place: its integer value v (of the digits, or the closing-factor int32) becomes
// rint(v 2^-k), ties to even, the value under E (re-split into balanced digits). v itself is
// rint(x s 2^k) for the scale s of E, so the result is within one unit of rint(x s).
// One workgroup per (matrix, 32-row strip, half): 32 rows x 256 columns, 8 planes (digits, 64
// pieces of 1 KB: block, 32-column chunk, plane) or 64 KB of closing-factor values; a lane keeps
// one row throughout (digits: lane l, row l / 2, bytes 16 (l % 2) of each plane's row; closing
// factors: row l % 32, the 16 bytes at (32 (l / 32) + row) 16 of each kilobyte).
__global__ void __launch_bounds__(256) lc_oz_fix_kernel(unsigned char* __restrict__ y, const int* __restrict__ eh,
                                                        int* __restrict__ ex, int P, int cf) {
  const int nst = P >> 5;
  const long long t = blockIdx.x;
  const int h = (int)(t & 1);
  const long long u = t >> 1;
  const int strip = (int)(u % (4 * nst));
  const long long m = u / (4 * nst);
  const int a = strip / nst, i0 = (strip - a * nst) * 32;
  const long long P2 = (long long)P * P;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int* ehm = eh + m * 8 * P;
  if (h == 0 && tid < 32) {
    const int r = a * P + i0 + tid;
    const int e0 = ehm[2 * r], e1 = ehm[2 * r + 1];
    const int E = e0 > e1 ? e0 : e1;
    ex[m * 4 * P + r] = E == INT_MIN ? 0 : E;
  }
  const int rr = cf ? (lane & 31) : (lane >> 1);
  const int r = a * P + i0 + rr;
  const int e0 = ehm[2 * r], e1 = ehm[2 * r + 1];
  const int eo = h ? e1 : e0, E = e0 > e1 ? e0 : e1;
  if (eo == INT_MIN || eo >= E) return;  // (no barrier in this kernel)
  const double sc = __builtin_ldexp(1.0, eo - E);
  unsigned char* yb = y + (m * 16 + a * 4 + 2 * h) * 8 * P2;  // block (a, 2 h)
  if (cf) {
    // 2 blocks x nst x nst chunks of 8 KB, those of the strip's rows: (cb, jc) at (jc nst + i0 / 32) 8 KB
    const int npc = 2 * nst * 8;
    for (int pc = wv; pc < npc; pc += 4) {
      const int cb = pc / (nst * 8), rest = pc - cb * nst * 8, jc = rest >> 3, k = rest & 7;
      i32x4* q = reinterpret_cast<i32x4*>(yb + cb * 8 * P2 + ((long long)jc * nst + (i0 >> 5)) * 8192 + k * 1024 +
                                          (32 * (lane >> 5) + rr) * 16);
      i32x4 v = *q;
#pragma unroll
      for (int e = 0; e < 4; e++) v[e] = lo32(__builtin_fma((double)v[e], sc, RND));
      *q = v;
    }
    return;
  }
  // digits: 2 blocks x nst chunks of 32 columns; per chunk 8 planes, P^2 bytes apart
  for (int pc = wv; pc < 2 * nst; pc += 4) {
    const int cb = pc / nst, jc = pc - cb * nst;
    unsigned char* q = yb + cb * 8 * P2 + ((long long)jc * P + i0 + rr) * 32 + 16 * (lane & 1);
    unsigned w[8][4];
#pragma unroll
    for (int pl = 0; pl < 8; pl++) {
      const i32x4 v = *reinterpret_cast<const i32x4*>(q + pl * P2);
#pragma unroll
      for (int g = 0; g < 4; g++) w[pl][g] = (unsigned)v[g];
    }
    unsigned o[8][4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      int dg[2][4][4];  // [re / im][column g 4 + b][digit]
#pragma unroll
      for (int part = 0; part < 2; part++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          int v = 0;
#pragma unroll
          for (int d = 0; d < 4; d++) v = v * 255 + (int)(signed char)(w[4 * part + d][g] >> (8 * b));
          digits_fast((double)v, sc, dg[part][b]);
        }
#pragma unroll
      for (int part = 0; part < 2; part++)
#pragma unroll
        for (int d = 0; d < 4; d++)
          o[4 * part + d][g] = pack4p(dg[part][0][d], dg[part][1][d], dg[part][2][d], dg[part][3][d]);
    }
#pragma unroll
    for (int pl = 0; pl < 8; pl++)
      *reinterpret_cast<i32x4*>(q + pl * P2) = i32x4{(int)o[pl][0], (int)o[pl][1], (int)o[pl][2], (int)o[pl][3]};
  }
}

// Digits back to FP64 (tests and diagnostics): D-space or elemental digit matrices, row-line
// planes, with one exponent per D-row (mode 1: block row R, row i at R * P + i, as a product's
// digit output has them) or per line index (mode 0, as the input conversion makes them).
__global__ void __launch_bounds__(256) lc_oz_decode_kernel(const unsigned char* __restrict__ y, const int* __restrict__ ex,
                                                           int nv, int P, int nblk, int mode, long long nmat,
                                                           double2* __restrict__ out) {
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nv2 = (long long)nv * nv;
  if (gid >= nmat * nblk * nv2) return;
  const long long bi = gid / nv2;
  const long long m = bi / nblk;
  const int blk = (int)(bi - m * nblk);
  const int row = (int)((gid / nv) % nv), col = (int)(gid % nv);
  const long long P2 = (long long)P * P;
  const unsigned char* b = y + bi * 8 * P2 + ((long long)(col >> 5) * P + row) * 32 + (col & 31);
  const int nr = mode ? 4 * P : P;
  const int E = ex[m * nr + (mode ? (blk >> 2) * P : 0) + row];
  double v[2];
  for (int part = 0; part < 2; part++) {
    long long s = 0;
    for (int d = 0; d < 4; d++) s = s * 255 + (signed char)b[(part * 4 + d) * P2];
    v[part] = __builtin_ldexp((double)s, E - 30) / SIGMA;
  }
  out[gid] = make_double2(v[0], v[1]);
}

}  // namespace lc

// NW = 4 for P = 32 (strips of 128 columns), 8 otherwise (256 columns); K256 for P <= 64.
// run: every group of the launch is a run (LcOzGroup::arun > 1; P <= 64). xtr: a group of the
// launch has more than LC_MAX_FUSED_TRACES traces or elemental closing factors. chain: a group of
// the launch is chained (LcOzGroup::pre; P <= 64, neither run nor xtr).
template <bool XTR, bool CHAIN>
static void launch_oz(dim3 grid, const LcOzGroup* d_groups, int ng, const LcOzGroup* d_twins, long long total_wg,
                      int nv, int P, int run, hipStream_t s) {
  if (P == 32 && run)
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<4, true, true, XTR, false>), grid, dim3(256), 0, s, d_groups, ng,
                       d_twins, nv, P, total_wg);
  else if (P == 32)
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<4, true, false, XTR, CHAIN>), grid, dim3(256), 0, s, d_groups, ng,
                       d_twins, nv, P, total_wg);
  else if (P == 64 && run)
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<8, true, true, XTR, false>), grid, dim3(512), 0, s, d_groups, ng,
                       d_twins, nv, P, total_wg);
  else if (P == 64)
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<8, true, false, XTR, CHAIN>), grid, dim3(512), 0, s, d_groups, ng,
                       d_twins, nv, P, total_wg);
  else
    hipLaunchKernelGGL((lc::lc_gemm_oz_kernel<8, false, false, XTR, false>), grid, dim3(512), 0, s, d_groups, ng,
                       d_twins, nv, P, total_wg);
}

extern "C" hipError_t lc_launch_gemm_oz(const LcOzGroup* d_groups, int ng, const LcOzGroup* d_twins,
                                        long long total_wg, int nv, int P, int run, int xtr, int chain,
                                        hipStream_t s) {
  if (ng <= 0 || total_wg <= 0) return hipSuccess;
  if (run && P > 64) return hipErrorInvalidValue;
  if (chain && (run || xtr || P > 64)) return hipErrorInvalidValue;
  // persistent: as many workgroups per XCD as are resident at once (one per CU), at most one per
  // logical workgroup
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 8)
      ncu = 256;
  }
  const long long slots = ncu / 8;  // (the stage buffers take 128 KB of LDS: one workgroup per CU)
  const long long per = (total_wg + 7) / 8;
  const dim3 grid((unsigned)(8 * (per < slots ? per : slots)));
  if (xtr) launch_oz<true, false>(grid, d_groups, ng, d_twins, total_wg, nv, P, run, s);
  else if (chain) launch_oz<false, true>(grid, d_groups, ng, d_twins, total_wg, nv, P, run, s);
  else launch_oz<false, false>(grid, d_groups, ng, d_twins, total_wg, nv, P, run, s);
  return hipGetLastError();
}

// x: nmat matrices of nblk blocks (nv x nv complex doubles), `stride` complex elements apart, to
// digits: y_rows / y_trans (nmat * nblk * 8 P^2 bytes each), ex_r / ex_t (nmat * P ints each).
// y_cf (optional, the size of y_rows): the closing-factor form under the row exponents; y_cft
// (optional): the closing-factor form of every transposed block under the column exponents.
extern "C" hipError_t lc_launch_oz_convert(const void* x, long long stride, long long nmat, int nv, int P, int nblk,
                                           int* ex_r, int* ex_t, void* y_rows, void* y_trans, void* y_cf, void* y_cft,
                                           hipStream_t s) {
  if (nmat <= 0) return hipSuccess;
  if (!y_rows || !y_trans || !ex_r || !ex_t || P > 128) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lc::lc_oz_exps_kernel, dim3((unsigned)nmat), dim3(256), 0, s, (const double2*)x, stride, nv, P,
                     nblk, ex_r, ex_t);
  const long long tiles = nmat * nblk * (long long)(P / 32) * (P / 32);
  hipLaunchKernelGGL(lc::lc_oz_digits_kernel, dim3((unsigned)tiles), dim3(256), 0, s, (const double2*)x, stride, nv,
                     P, nblk, (const int*)ex_r, (const int*)ex_t, (unsigned char*)y_rows, (unsigned char*)y_trans,
                     (int2*)y_cf, (int2*)y_cft);
  return hipGetLastError();
}

// nmat consecutive D-space digit matrices (8 * 16 P^2 bytes each; closing-factor form with cf)
// written by a P = 128 product with half-row exponents eh (8 P ints per matrix): the D-row
// exponents into ex (4 P ints per matrix) and each row's halves under them (lc_oz_fix_kernel).
extern "C" hipError_t lc_launch_oz_fix(void* y, const int* eh, int* ex, long long nmat, int P, int cf, hipStream_t s) {
  if (nmat <= 0) return hipSuccess;
  if (P != 128) return hipErrorInvalidValue;  // (only there is a strip half a D-row)
  const long long wgs = nmat * 4 * (P / 32) * 2;
  if (wgs > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lc::lc_oz_fix_kernel, dim3((unsigned)wgs), dim3(256), 0, s, (unsigned char*)y, eh, ex, P, cf);
  return hipGetLastError();
}

extern "C" hipError_t lc_launch_oz_decode(const void* y, const int* ex, int nv, int P, int nblk, int mode,
                                          long long nmat, void* out, hipStream_t s) {
  const long long n = nmat * nblk * (long long)nv * nv;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(lc::lc_oz_decode_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const unsigned char*)y, ex, nv, P, nblk, mode, nmat, (double2*)out);
  return hipGetLastError();
}
