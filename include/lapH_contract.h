/* C-ABI of the MI355X LapH meson contraction library (libcontract_hip.so).
 *
 * Plain pointers and sizes only. Every function returns 0 on success or a negative LC_ERR_*
 * code; lc_last_error() returns the message of the last failure. "d_" pointers are device
 * memory (hipMalloc / torch), "h_" pointers host memory. Complex numbers are interleaved
 * (re, im) of the compute type: prec = LC_SINGLE -> float, LC_DOUBLE -> double, the reference's
 * PRECISION enum (include/contract.h:33).
 *
 * Input layouts are the reference's device layouts (lib/contract.cpp:171-219):
 *   row   tau(tf, .)          [nt][4][4][nv][nv]   one sink row of the perambulator
 *   diag  tau(t, t)           [nt][4][4][nv][nv]   (lib/contract.cpp:549-570)
 *   mom5  Mom_matrices        [nt][5][nv][nv]
 *   x2    Mom_matrices_x2     [nt][2][nv][nv]
 *   mom, p1, p2, m1..m4       [nt][nv][nv]
 * Outputs land in host arrays with the layouts the reference launchers write into h_elem.
 */
#ifndef LAPH_CONTRACT_H
#define LAPH_CONTRACT_H

#ifdef __cplusplus
extern "C" {
#endif

enum { LC_SINGLE = 0, LC_DOUBLE = 1 };
enum { LC_I0 = 0, LC_I1 = 1 };

enum {
  LC_OK = 0,
  LC_ERR_NOT_INIT = -1,   /* library not initialised (lib/contract.cpp:168)            */
  LC_ERR_ARG = -2,        /* bad size, precision, index or null pointer                 */
  LC_ERR_HIP = -3,        /* HIP runtime error (message carries hipGetErrorString)      */
  LC_ERR_ALREADY = -4,    /* initialised twice (lib/contract.cpp:62-63)                 */
  LC_ERR_NODEV = -5,      /* no GPU or bad device id (lib/contract.cpp:33-49)           */
  LC_ERR_NOMEM = -6,      /* device allocation failed                                   */
  LC_ERR_IO = -7          /* file missing, header mismatch or short read (:107-130)     */
};

/* Library state. Replaces initContract/initDevice/destroyContract (lib/contract.cpp:25-88)
 * for the kernel layer: selects the device, creates the stream, owns persistent scratch. */
int lc_init(int device);
int lc_finalize(void);
int lc_is_init(void);
const char* lc_last_error(void);
int lc_device_count(void);

/* Sink-row launchers. Each replaces one reference run_* function (include/contract.h:50-69)
 * with device pointers in place of texture objects, and writes the spin-summed,
 * diagram-weighted per-source values of one sink row tf into host memory. */

/* run_ContractRho (lib/contract_kernels_rho.cu:126-138): h_corr [nt][16][3] complex. */
int lc_rho_row(int prec, const void* d_row, const void* d_mom5, int tf, int nt, int nv,
               void* h_corr);

/* run_ContractRhoPi (lib/contract_kernels_rhopi.cu:125-137): h_corr [nt][4] complex. */
int lc_rhopi_row(int prec, const void* d_row, const void* d_diag, const void* d_p1,
                 const void* d_p2, const void* d_p12x5, int tf, int nt, int nv, int idir,
                 void* h_corr);

/* run_ContractPiPi / run_ContractPiPi_I0 (lib/contract_kernels_pipi.cu:507-533):
 * iso = LC_I1 or LC_I0; h_corr [nt][5] complex = weighted {square, doubleTriangle, star, fish,
 * doubleTriangleHor (I0 only, else 0)} (lib/contract_kernels_pipi.cu:288-303, 463-478). */
int lc_pipi_row(int prec, int iso, const void* d_row, const void* d_diag, const void* d_m1,
                const void* d_m2, const void* d_m3, const void* d_m4, int tf, int nt, int nv,
                void* h_corr);

/* run_ContractLoopUnity (lib/contract_kernels_loop_unity.cu:95-107): h_loop [nt] complex. */
int lc_loop_unity(int prec, const void* d_diag, const void* d_mom, int nt, int nv, void* h_loop);

/* lc_loop_unity without any allocation, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops); no counterpart in the reference. Same inputs; the result
 * d_loop [nt] complex (prec) is written on the device, ordered on the library stream (no
 * synchronisation: lc_synchronize or a later library call). Its partial sums use one fixed device
 * array of 65536 complex doubles (1 MiB, held on each initialised device until the process ends):
 * Nt * max(1, min(64, ceil(4 NeV^2 / 4096))) must not exceed 65536, else LC_ERR_ARG. Calls that
 * overlap on one device share that array and race. */
int lc_loop_unity_noalloc(int prec, const void* d_diag, const void* d_mom, int nt, int nv, void* d_loop);

/* run_ContractSigma (lib/contract_kernels_sigma.cu:134-147): h_corr [nt][4] complex. */
int lc_sigma_row(int prec, const void* d_row, const void* d_x2, int tf, int nt, int nv,
                 void* h_corr);

/* run_ContractSigmaPi_fish_hor (lib/contract_kernels_sigmapi.cu:190-202): h_corr [nt]. */
int lc_sigmapi_fish_hor(int prec, const void* d_diag, const void* d_p1, const void* d_p2,
                        int nt, int nv, void* h_corr);

/* run_ContractSigmaPi_triangle (lib/contract_kernels_sigmapi.cu:204-217): h_corr [nt][2]. */
int lc_sigmapi_triangle_row(int prec, const void* d_row, const void* d_diag, const void* d_p1,
                            const void* d_p2, const void* d_x2, int tf, int nt, int nv,
                            void* h_corr);

/* Batched engine (no reference counterpart; the batching form that replaces the reference's
 * per-sink-row loops, lib/contract.cpp:602-605 and :775-779, for a resident perambulator):
 * the rho (op pair (0,0), Gamma = gamma_i) + pi-pi (back-to-back Phi(p), Phi(p)^dag) set for
 * every (sink tf, source t0 = src0 + k*src_step, momentum p), FP64.
 *   d_tau [nt][nt][4][4][nv][nv], d_phi [np][nt][nv][nv]
 *   d_raw  [np][nsrc][nt][10]  {rho_0..2, square, dT, dTHor, star1, star2, fish1, fish2} traces
 *   d_corr [np][nsrc][nt][8]   {rho_0..2, square, dT, dTHor, star, fish} correlators
 *   d_avg  [np][8][nt]         source-averaged C(dt), FP64 on the device
 * Unweighted diagrams; max_scratch_bytes bounds the intermediate storage (<= 0: 60% of the
 * device memory free at the call). */
int lc_engine_rho_pipi(const void* d_tau, const void* d_phi, int nt, int nv, int np, int src0,
                       int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                       void* d_corr, void* d_avg);

/* Mixed-precision form of lc_engine_rho_pipi, same arguments and outputs. Two forms
 * (lc_set_mixed_form):
 *  LC_MIXED_EXACT (default): every real value of an operand is held as four balanced int8
 *    digits of base 255 under a power-of-two exponent per matrix line (about 31 bits below the
 *    line's largest magnitude); products run on the int8 matrix cores with exact int32 accumulation of
 *    the digit pairs of total order <= 3, and every product, closing trace and source sum is
 *    FP64 (nv <= 128);
 *  LC_MIXED_BF16X3: tau and Phi rescaled per time slice to unit Frobenius norm (scales kept in
 *    FP64) and stored as FP32; products on the BF16 matrix cores with an exact three-way BF16
 *    split of each FP32 operand (six cross products, FP32 accumulation); FP64 traces. */
#define LC_MIXED_EXACT 0
#define LC_MIXED_BF16X3 1
int lc_set_mixed_form(int form);
int lc_engine_rho_pipi_mixed(const void* d_tau, const void* d_phi, int nt, int nv, int np, int src0,
                             int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                             void* d_corr, void* d_avg);

/* Batched pi-pi over explicit momentum sets: the engine of lc_engine_rho_pipi for arbitrary
 * elementals P1..P4 per set, the four Mom_matrix arguments of
 * Contract_pipi_pipi::contractPiPi (lib/contract.cpp:748-779; the driver's momentum
 * combinations px1..px4, tests/compute_pipi_I0.cpp:130-150), batched over every set, sink and
 * source. P1, P2 sit at the source t0 and P3, P4 at the sink tf, as in run_ContractPiPi
 * (lib/contract_kernels_pipi.cu:521).
 *   d_elem [nset][4][nt][nv][nv]   {P1, P2, P3, P4} of each set, complex double
 *   d_raw / d_corr / d_avg         as lc_engine_rho_pipi, with np = nset
 * The pi-pi slots (square, dT, dTHor, star, fish) are the unweighted diagrams of the set; the
 * rho slots hold -Tr[(P4(tf) tau)^dag (g5 g_i) (tau P1(t0)) (-g_i g5)]. */
int lc_engine_pipi_sets(const void* d_tau, const void* d_elem, int nt, int nv, int nset, int src0,
                        int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_corr,
                        void* d_avg);

/* Mixed-precision form of lc_engine_pipi_sets (as lc_engine_rho_pipi_mixed). */
int lc_engine_pipi_sets_mixed(const void* d_tau, const void* d_elem, int nt, int nv, int nset, int src0,
                              int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                              void* d_corr, void* d_avg);

/* Batched sigma / sigma-pi / rho-pi set (no reference counterpart: the reference forms these one
 * sink row at a time through run_ContractSigma, run_ContractSigmaPi_triangle, run_ContractRhoPi,
 * run_ContractLoopUnity and run_ContractSigmaPi_fish_hor). On a resident perambulator, for every
 * momentum set p, sink tf and source t0_k = src0 + k * src_step:
 *   d_tau   [nt][nt][4][4][nv][nv]   complex double
 *   d_el    [np][nt][11][nv][nv]     complex double elementals by type: 0-1 sigma X_l, 2-3 sigma-pi
 *                                    triangle X_ip, 4-8 rho-pi P1P2 types, 9 P1, 10 P2
 *   d_raw   [np][nsrc][nt][10]       {sigma(l,r) l*2+r, triangle(0..1), rho-pi(iop 0..3)}: the
 *                                    unweighted per-source values of the row launchers
 *   d_loops [np][nt][3]              {Tr[D(t) X_0(t)], Tr[D(t) X_1(t)], hor_fish(t)}
 * idir: the rho polarisation of the rho-pi operators. */
int lc_engine_sigma_set(const void* d_tau, const void* d_el, int nt, int nv, int np, int idir, int src0,
                        int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_loops);

/* Mixed-precision form of lc_engine_sigma_set (same arguments and outputs), in the form
 * lc_set_mixed_form selects: LC_MIXED_EXACT (default; int8 digits with exact integer
 * accumulation, FP64 products and traces, nv <= 128) or LC_MIXED_BF16X3 (per-slice rescaled FP32
 * values split into BF16 pieces on the BF16 matrix cores, FP64 traces and scale factors). The
 * per-time loops and hor_fish are FP64 in both. */
int lc_engine_sigma_set_mixed(const void* d_tau, const void* d_el, int nt, int nv, int np, int idir,
                              int src0, int src_step, int nsrc, long long max_scratch_bytes, void* d_raw,
                              void* d_loops);

/* hipGraph replay of the batched engines (no reference counterpart). enable = 1: the first
 * lc_engine_* call with a given set of arguments (pointers, sizes, source list, scratch budget) and
 * tile parameters runs eagerly and is recorded; later calls with the same arguments replay the
 * recorded launches as one graph (the inputs are read again, so new values in the same buffers are
 * used). A call is recorded again after the library's shared scratch or partial buffers were
 * reallocated. Profiled calls (lc_profile) and the autotuner's candidate runs are never recorded.
 * enable = 0 destroys the graphs. Needs no lc_init. */
int lc_engine_graph(int enable);
/* What the last engine call did: *state = 0 eager, 1 recorded, 2 replayed. */
int lc_engine_graph_report(int* state);

/* doubleTriangleHor with the conjugate transpose tau^dag (no reference counterpart). The reference's
 * kernel uses W_ab = conj(tau_ba) element-wise, without transposing the eigenvector indices
 * (lib/calculate_pipi_doubleTriangleHor_core.h:126); that stays the default (on = 0) of lc_pipi_row,
 * calculatePiPi and the rho + pi-pi engines. on = 1 switches all of them to W = tau^dag. Needs no
 * lc_init. */
int lc_set_dthor_dagger(int on);

/* Tile shape of the mixed-precision GEMM launches (no reference counterpart): 64 x 64 output tiles
 * per workgroup, 1, 2 or 4; 0 restores the built-in rule (4 when Nv <= 64, else 1). Applies to the
 * *_mixed engines (the FP64 kernels take one tile per workgroup) and takes precedence over the
 * autotuner. Results do not depend on it. Needs no lc_init. */
int lc_set_tile_params(int tiles_per_wg);

/* In-project autotuner (no reference counterpart). enable = 1: the first *_mixed engine call of
 * each shape (engine, nt, nv, np, idir, source list, momentum chunk) runs itself for every
 * candidate tile shape (4, 2 and 1 tiles per workgroup, twice each, timed with HIP events; the
 * engines overwrite their outputs, so the repeats give the same results), keeps the fastest for
 * the later calls of that shape and leaves the fastest's results in the outputs. enable = 0 stops
 * tuning and uses the built-in rule (choices are kept); -1 also forgets them. lc_finalize forgets
 * them too. Needs no lc_init. */
int lc_autotune(int enable);

/* What the last engine call ran with: out5 = {tiles per workgroup (0: built-in rule), how they
 * were chosen (0 fixed, 1 cached choice, 2 tuned by that call), ms of the candidates 4, 2 and 1
 * (best of two runs; 0 unless tuned by that call)}. */
int lc_autotune_report(double* out5);

/* Perambulator file -> device: the reference's readPropH (lib/contract.cpp:91-160) with the file
 * order [s1][v1][t1][s2][v2][t2] permuted on the GPU into d_tau [nt][nt][4][4][nv][nv] (complex
 * double, or complex float for prec = LC_SINGLE), streamed one (s1, v1) slab at a time through
 * pinned staging buffers. LC_ERR_IO on a missing file, a header that does not match (nt, nv) or a
 * short read. */
int lc_read_prop_device(const char* path, int nt, int nv, int prec, void* d_tau);

/* Plain momentum elementals from Laplacian eigenvectors, on the GPU: replaces the drivers' host loop
 * matp[t](r,c) = cscalar_product(evec[r], phase * evec[c]) (tests/compute_rho.cpp:116-129):
 *   d_phi[p][t][r][c] = sum_k conj(d_evec[t][r][k]) * d_phase[p][k] * d_evec[t][c][k]
 * d_evec [nt][nv][nsite] complex double (k over site and colour, in the order the phases follow),
 * d_phase [np][nsite] complex double (exp(-2 pi i p.x / L) at the site of entry k),
 * d_phi [np][nt][nv][nv] complex double: the type-0 layout of the contraction inputs. FP64. */
int lc_build_elementals(const void* d_evec, const void* d_phase, int nt, int nv, long long nsite, int np,
                        void* d_phi);

/* Covariant derivatives of the Laplacian eigenvectors, the drivers' lm.covderiv(j, evec, dl[j])
 * (tests/compute_rho.cpp:140-141, from the external gwu-qcd library), on the GPU:
 *   d_dvec[j][t][c][3x + a] = c [U_j(x) v(x + j) - U_j(x - j)^dag v(x - j)]_a,  v = d_evec[t][c],
 * periodic in the spatial directions j = 0, 1, 2 (x, y, z). Sites x = ix + nx (iy + ny iz), colour
 * fastest: d_evec [nt][nv][3 nx ny nz], d_links [nt][3][nx ny nz][3][3] (row-major SU(3) U_j(x)),
 * d_dvec [3][nt][nv][3 nx ny nz], complex double. The normalisation c of gwu-qcd's covderiv stencil
 * is not in the reference (an external library): lc_covariant_derivative uses the central
 * difference, c = 1/2; lc_covariant_derivative_scaled takes c. Type 1 of lc_build_elementals_types
 * scales as c^2, types 2-4 as c. */
int lc_covariant_derivative(const void* d_links, const void* d_evec, int nt, int nv, int nx, int ny, int nz,
                            void* d_dvec);
int lc_covariant_derivative_scaled(const void* d_links, const void* d_evec, int nt, int nv, int nx, int ny, int nz,
                                   double c, void* d_dvec);

/* The five elemental types of the rho / rho-pi drivers (tests/compute_rho.cpp:116-229, the tmpMom
 * layout) for np momenta, from eigenvectors and their covariant derivatives:
 *   type 0      sum_k conj(V[r][k]) ph[k] V[c][k]                         (plain)
 *   type 1      sum_j sum_k conj(dV_j[r][k]) ph[k] dV_j[c][k]              (smeared)
 *   type 2 + j  sum_k conj(V[r][k]) ph[k] dV_j[c][k] - conj(dV_j[r][k]) ph[k] V[c][k]
 * d_evec [nt][nv][nsite], d_dvec [3][nt][nv][nsite] (lc_covariant_derivative), d_phase [np][nsite],
 * d_mom5 [np][5][nt][nv][nv], complex double, FP64 matrix cores. */
int lc_build_elementals_types(const void* d_evec, const void* d_dvec, const void* d_phase, int nt, int nv,
                              long long nsite, int np, void* d_mom5);

/* Diagnostic: one v_mfma_f32_16x16x32_bf16 on A(16x32), B(32x16) given as FP32 values exact in
 * BF16; h_d as for lc_selftest_mfma (D row-major, then raw accumulators). */
int lc_selftest_mfma_bf16(const float* h_a, const float* h_b, float* h_d);

/* Kernel timing with HIP events on the library stream: lc_profile(1) clears and enables,
 * lc_profile_read fills {gemm_ms, gemm_launches, gemm_flops, trace_ms, trace_launches,
 * trace_flops} accumulated since. */
int lc_profile(int enable);
int lc_profile_read(double* out6);

/* Whole-call drop-ins for FFI callers: the reference's initContract / destroyContract /
 * readPropH / calculate* flows (lib/contract.cpp:58-160, lib/interface.cpp:393-426) with error
 * codes in place of exit(). prop is the host buffer of readPropH ([nt][nt][4][4][nev][nev] of
 * the precision's type), elementals are FP64 driver arrays [type][nt][nev][nev]; the text output
 * is appended to outfile in the reference's format. lc_set_binary_output(1) also appends the
 * printed values as raw doubles to <outfile>.bin. */
int lc_contract_init(int nev, int nt, int dev);
int lc_contract_destroy(void);
int lc_set_binary_output(int on);
int lc_set_i_dir(int idir);  /* the process-global direction the rho-pi text prints */
int lc_read_prop(const char* path, int prec, void** out);
void lc_free_host(void* p);
int lc_calculate_rho(int nev, int nt, void* prop, double* mom, int prec, const char* outfile);
int lc_calculate_rhopi(int idir, int nev, int nt, void* prop, double* p1, double* p2, double* p12,
                       int prec, const char* outfile);
int lc_calculate_pipi(int nev, int nt, void* prop, double* p1, double* p2, double* p3, double* p4,
                      int prec, const char* outfile, int iso);
int lc_calculate_sigma(int nev, int nt, void* prop, double* mom, int prec, const char* outfile);
int lc_calculate_sigmapi(int nev, int nt, void* prop, double* p1, double* p2, double* p12, int prec,
                         const char* outfile);

/* Device memory on the library's own HIP runtime and stream (for callers without a GPU
 * framework; no reference counterpart: the reference allocated inside each Base object). */
int lc_dev_alloc(long long bytes, void** out);
int lc_dev_free(void* p);
int lc_memcpy_h2d(void* d_dst, const void* h_src, long long bytes);
int lc_memcpy_d2h(void* h_dst, const void* d_src, long long bytes);
int lc_memcpy_d2d(void* d_dst, const void* d_src, long long bytes);
int lc_synchronize(void);

/* Diagnostic: one 16x16x4 matrix-core product A(16x4)*B(4x16). h_d[0..255] receives D
 * row-major as the kernels map accumulators, h_d[256..511] the raw accumulator element q of
 * lane l at 256 + 4l + q. Used by the tests to pin the layout. No reference counterpart. */
int lc_selftest_mfma(int prec, const void* h_a, const void* h_b, void* h_d);

/* Diagnostic of the exact form (int8 digit storage, exact integer accumulation): one product
 * C = A' B' of host FP64 operands, D-space [16][nv][nv] or elemental [nv][nv] (ka / kb = 1),
 * A' = G(gam4[0]) f_fa(A) G(gam4[1]), B' = G(gam4[2]) f_fb(B) G(gam4[3]) with f a combination of
 * LC_VIEW_CONJ | LC_VIEW_TV | LC_VIEW_TS (element conjugate, eigenvector transpose, spin-block
 * transpose) and G(mu) = gamma_mu (0: none; D-space operands only). h_c receives C
 * [16][nv][nv] in FP64; h_cd (optional) C after its round trip through the product's own digit
 * output (at nv > 64 with half-row exponents and the fix-up that makes them whole rows). nv <= 128.
 * No reference counterpart. */
#define LC_VIEW_CONJ 1
#define LC_VIEW_TV 2
#define LC_VIEW_TS 4
int lc_selftest_oz_gemm(int nv, int ka, int kb, int fa, int fb, const int* gam4, const void* h_a,
                        const void* h_b, void* h_c, void* h_cd);
/* The same product with a closing-factor output (one int32 per real, v = rint(x 2^(30 - E)) under
 * the D-row exponent E; at nv > 64 written under half-row exponents and brought to E by the
 * fix-up) returned as stored: h_cf receives 16 blocks of 8 P^2 bytes (P = 32, 64 or 128 for
 * nv <= 32, 64, 128), element (r, j) of block b at 8 b P^2 + ((j/32) (P/32) + r/32) 8192 +
 * (2 ((j%32)/8) + (j/2)%2) 1024 + (32 ((j/4)%2) + r%32) 16 + 8 (j%2) bytes (re, im); h_e the 4 P
 * D-row exponents (block row R, row i at R P + i). No reference counterpart. */
int lc_selftest_oz_gemm_cf(int nv, int ka, int kb, int fa, int fb, const int* gam4, const void* h_a,
                           const void* h_b, void* h_c, void* h_cf, int* h_e);

#ifdef __cplusplus
}
#endif
#endif
