// Internal descriptors shared by the host planner and the HIP kernels.
//
// Every diagram is evaluated in spin (x) eigenvector space ("D-space", D = 4*Nv). A D-space
// matrix is stored as 4x4 spin blocks of Nv x Nv complex entries, block (a,b) at
// (a*4+b)*Nv*Nv -- the perambulator block layout [s1][s2][v1][v2] of the reference
// (lib/contract.cpp:140,178). Elementals are single Nv x Nv blocks acting as M (x) 1_spin.
//
// Gamma structure never becomes arithmetic: a spin matrix with one non-zero per row and
// column (gamma_5, gamma_i, gamma_5 gamma_i, ...) multiplying a D-space matrix only renames
// blocks and multiplies them by an exact phase (+-1, +-i, +-1/2). An LcView carries those
// renames and phases together with the element operation (conj, transpose of the eigenvector
// indices, transpose of the spin-block indices), so a kernel reading block (a,b) of a view reads
//
//    ph_r[a] * ph_c[b] * f( X[ perm_r[a] ][ perm_c[b] ] )      (spin indices swapped if LC_TS)
//
// where f applies LC_CONJ / LC_TV element-wise.
#pragma once
#include <stdint.h>

enum { LC_DENSE = 0, LC_KRON = 1 };
enum { LC_CONJ = 1, LC_TV = 2, LC_TS = 4 };
// X^dagger of a D-space matrix: conj + eigenvector transpose + spin-block transpose.
#define LC_DAG (LC_CONJ | LC_TV | LC_TS)

struct LcView {
  const void* base;         // complex elements, interleaved (re, im)
  long long s[3];           // batch strides, in complex elements
  int kind;                 // LC_DENSE or LC_KRON
  int flags;                // LC_CONJ | LC_TV | LC_TS
  int perm_r[4], perm_c[4];
  double ph_r[4][2], ph_c[4][2];
  // digit storage only (exact form, gemm_oz.hip): the power-of-two exponents of the stored lines,
  // line i of stored block row R at ex[voff(es) + R * ebs + i] (ebs = 0: one exponent per line
  // index shared by all blocks of the matrix)
  const int* ex;
  long long es[3];
  int ebs;
};

#define LC_MAX_FUSED_TRACES 3
// fused traces of one product of the exact kernel (gemm_oz.hip), whose closing factors may also be
// elemental views with a spin matrix (see LcOzGroup)
#define LC_OZ_MAX_TRACES 6

// C = alpha * A * B over a batch (i0 < n0, i1 < n1, i2 < n2). C is a plain D-space matrix
// (not stored when C == nullptr). Optionally fused closing traces: for t < ntr,
//   part_t[item][tile] = tr_alpha_t * sum over the tile of C_ij * Y_t(c,a)_ji
// i.e. the per-tile pieces of tr_alpha_t * Tr[C * Y_t]; a fixed-order reduction sums the tiles.
//
// Split storage (mixed-precision kernel only). A split matrix holds every complex value x as six
// BF16 planes [piece][re/im]: x = x0 + x1 + x2 exactly for each component (three-way split of an
// FP32 value, see gemm_mx.hip). Block (a,b) of a D-space matrix, or the single block of an
// elemental, occupies 6 * Nv * Nv BF16 values, plane q = 2 * piece + (0 re, 1 im) at q * Nv * Nv,
// each plane Nv x Nv row-major. Offsets and strides stay in complex elements (6 BF16 each), so a
// split view is addressed exactly like a complex view of the same shape. The mixed kernel reads
// its A and B operands in split storage; C may be written as FP32 complex (C) and/or split (Cs,
// every block transposed when cs_trans: element (i,j) at j * Nv + i).
struct LcGemmGroup {
  LcView A, B;
  LcView B2;                // sandwich groups (mixed-precision kernel): C = alpha * A B B2
  // optional second right factor of a sandwich (B3.base != nullptr): the product A B B3 is not
  // stored, only traced, trx_alpha * Tr[(A B B3) Yx], into partx
  LcView B3, Yx;
  double trx_alpha[2];
  double* partx;
  void* C;
  void* Cs;
  int cs_trans;
  int bpw;                  // output tiles per workgroup (mixed-precision kernel)
  int col_tiles;            // bpw tiles of a workgroup: 0 one block row, 1 one block column
  long long cs[3];
  double alpha[2];
  int n[3];
  int tiles;                // output tiles per block side: ceil(nv / TM)
  int ntr;
  long long wg_begin;       // first workgroup of this group in the launch
  LcView Y[LC_MAX_FUSED_TRACES];
  double tr_alpha[LC_MAX_FUSED_TRACES][2];
  double* part[LC_MAX_FUSED_TRACES];  // complex double [items][16*tiles*tiles]
};

// out[item] (+)= sum_{k < nk} part[item*nk + k] (fixed order), complex conjugated first when
// conj; out addressed through strides.
struct LcReduceJob {
  const double* part;
  void* out;                // complex of the output type
  long long os[3];
  int n[3];
  int nk;
  int accumulate;
  int out_double;           // 1: out is complex double, 0: complex float
  int conj;
};

// out (+)= alpha * Tr[X * Y] over a batch. out holds complex values of the compute type.
struct LcTraceGroup {
  LcView X, Y;
  void* out;
  long long os[3];
  double alpha[2];
  int n[3];
  int accumulate;
  long long wg_begin;
};

// Exact form (gemm_oz.hip). A digit matrix holds every real value x of a stored line as four
// balanced base-255 digits (int8, each in [-127, 127]) under the line's exponent E:
//   x = (D0 * 255^3 + D1 * 255^2 + D2 * 255 + D3) / (SIGMA * 2^(30 - E)),   |x| < 2^E,
// with SIGMA = 63/32, i.e. about 31 bits below the largest magnitude of the line. Blocks are
// P x P (P = Nv rounded up to a multiple of 32, zero padded); block (a,b) of a D-space matrix sits
// at (a*4+b) * 8 * P^2 bytes, plane q = part * 4 + digit (part 0 real, 1 imaginary) at q * P^2
// bytes of the block, and inside a plane element (line, k) at ((k / 32) * P + line) * 32 + k % 32:
// each 32-line x 32-K chunk is one contiguous kilobyte. Offsets and strides count 8-byte "complex
// digit elements" (P^2 per block), so a digit view is addressed like a complex-float view of P x P
// blocks.
//
// C = alpha * A * B over a batch with A, B digit views (A read by rows, B by columns, as in the
// mixed form). Digit products of total order <= 3 accumulate exactly in int32 (one accumulator
// per order), and the FP64 epilogue combines them, so the only rounding is the digit truncation
// of the operands and the dropped orders 4..6 (about 2^-32 of the largest products). Outputs:
// FP64 complex C (blocks of nv x nv), and/or digits Cd with one exponent per D-row written to Ce
// ([4][P] ints per matrix; at P = 128, where a workgroup strip is half a D-row, one per half,
// [4 P][2] ints, each half under its own, for lc_oz_fix_kernel), and fused closing traces as in
// the other kernels. A closing factor in the closing-factor form may be
// an elemental view with a spin matrix S (X (x) S, block (a, b) = ph_r[a] ph_c[b] X where
// perm_r[a] == perm_c[b], else zero); every other reader of an elemental view takes it as X (x) 1
// and requires the identity perms and phases.
struct LcOzGroup {
  LcView A, B;
  void* C;
  long long cs[3];
  void* Cd;
  int* Ce;
  int cf;                   // Cd in the closing-factor form instead of digits
  long long cds[3];         // in complex digit elements
  long long ces[3];         // in ints
  double alpha[2];
  int n[3];
  int kind;                 // 0 dense x dense, 1 elemental x dense, 2 dense x elemental
  int ntr;
  int nseg;                 // column segments of the workgroup strip per D-row
  // index of the group's twin in the launch's twin array, or -1. A twin is a second product with
  // the same A operand, kind and items and its own B and fused traces (no C or Cd), formed per item
  // right after this one from the A operand already staged on the chip (every K step of it: an
  // elemental factor, K <= 2 * 32 per block, or at P = 128 an elemental B, 4 steps of 2 chunks)
  int twin;
  // items per logical workgroup (1, or n[0] when A does not depend on axis 0): a workgroup stages
  // its strip of A once (every K step in its own stage buffer, K <= 8 * 32) and forms the products
  // of the n[0] items along axis 0 with it
  int arun;
  // index in the launch's side array (the one holding twins) of a first product whose output is
  // this product's A, or -1. Chained: per item and strip the kernel forms the first product
  // (elemental x dense), writes its digits straight into the stage buffers as this product's A
  // (dense x elemental; 4 P <= 256), and never stores it; this group's A view is unused
  int pre;
  long long wg_begin;
  LcView Y[LC_OZ_MAX_TRACES];
  double tr_alpha[LC_OZ_MAX_TRACES][2];
  double* part[LC_OZ_MAX_TRACES];
};
