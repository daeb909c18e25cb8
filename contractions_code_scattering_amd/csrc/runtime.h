// Host runtime of the contraction library: device context, persistent buffers, and the
// "program" builder that turns D-space chains into grouped GEMM / trace launches.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "lc_internal.h"
#include "../../include/lapH_contract.h"

extern "C" hipError_t lc_launch_gemm(int is_double, const LcGemmGroup* d_groups, int ng,
                                     long long total_wg, int nv, hipStream_t s);
extern "C" hipError_t lc_launch_trace(int is_double, const LcTraceGroup* d_groups, int ng,
                                      long long total_wg, int nv, hipStream_t s);
extern "C" hipError_t lc_launch_reduce(const LcReduceJob* d_jobs, const long long* d_begins,
                                       int njobs, long long total, hipStream_t s);
extern "C" hipError_t lc_launch_gemm_mx(int variant, const LcGemmGroup* d_groups, int ng,
                                        long long total_wg, int nv, hipStream_t s);
extern "C" hipError_t lc_launch_mfma_bf16_selftest(const float* a, const float* b, float* d,
                                                   hipStream_t s);
extern "C" hipError_t lc_launch_split_rescale(const void* x, void* y_rows, void* y_trans, void* y32,
                                              long long len, int nv, long long nslice, double* scale,
                                              hipStream_t s);
extern "C" hipError_t lc_launch_mfma_selftest(int is_double, const void* a, const void* b,
                                              void* d, hipStream_t s);
extern "C" int lc_gemm_tile(int nv);
extern "C" hipError_t lc_launch_gemm_oz(const LcOzGroup* d_groups, int ng, const LcOzGroup* d_twins,
                                        long long total_wg, int nv, int P, int run, int xtr, int chain,
                                        hipStream_t s);
extern "C" hipError_t lc_launch_oz_convert(const void* x, long long stride, long long nmat, int nv, int P, int nblk,
                                           int* ex_r, int* ex_t, void* y_rows, void* y_trans, void* y_cf, void* y_cft,
                                           hipStream_t s);
extern "C" hipError_t lc_launch_oz_fix(void* y, const int* eh, int* ex, long long nmat, int P, int cf, hipStream_t s);
extern "C" hipError_t lc_launch_oz_decode(const void* y, const int* ex, int nv, int P, int nblk, int mode,
                                          long long nmat, void* out, hipStream_t s);

namespace lc {

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error{code, buf};
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(LC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Grow-only device buffer. `gen` counts its allocations, so that a recorded graph can tell
// whether the memory it launches on is still the same.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  long long gen = 0;
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    gen++;
    if (p) {
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      hip_check(hipFree(p), "hipFree");
      p = nullptr;
      cap = 0;
    }
    size_t n = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&p, n) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      fail(LC_ERR_NOMEM, "hipMalloc of %zu bytes failed", n);
    }
    cap = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Per-launch timing with HIP events on the library stream (enabled by lc_profile(1)).
struct KernelTimes {
  double gemm_ms = 0, trace_ms = 0, other_ms = 0;
  long long gemm_launches = 0, trace_launches = 0;
  double gemm_flops = 0, trace_flops = 0;
  void clear() { *this = KernelTimes(); }
};

// What the last engine call ran with: output tiles per workgroup of the mixed GEMM (0: the
// built-in rule), how they were chosen, and the candidate timings when this call tuned them.
struct TuneRecord {
  int bpw = 0;
  int how = 0;  // 0 fixed (rule or lc_set_tile_params), 1 cached choice, 2 tuned by this call
  double ms[3] = {0, 0, 0};  // candidates 4, 2, 1 tiles per workgroup (best of two runs each)
};

// Launches of an engine call being recorded for graph replay, and the descriptor buffers they read.
struct LaunchLog {
  std::vector<std::function<hipError_t(hipStream_t)>> launches;
  std::vector<DevBuf> bufs;
};

// A recorded engine call: the instantiated graph, its descriptor buffers and the allocations of
// the shared scratch and partial buffers it was recorded on.
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  std::vector<DevBuf> bufs;
  long long scratch_gen = 0, partials_gen = 0;
};

struct Context {
  bool on = false;
  int device = -1;
  hipStream_t stream = nullptr;
  DevBuf groups, scratch, out, partials;
  bool profile = false;
  KernelTimes times;
  std::vector<hipEvent_t> ev_pool;
  // mixed-GEMM tiles per workgroup: the caller's override (0: rule), the value in force while an
  // engine call plans its launches, and the autotuner's switch, choices and last record
  // doubleTriangleHor with tau^dag (conjugate transpose) in place of the reference's element-wise
  // conjugate (lc_set_dthor_dagger)
  bool dthor_dagger = false;
  // form of the mixed-precision engines (lc_set_mixed_form): 0 exact (int8 digits, exact integer
  // accumulation, gemm_oz.hip), 1 BF16x3 split with FP32 accumulation (gemm_mx.hip)
  int mixed_form = 0;
  int tile_bpw = 0;
  int active_bpw = 0;
  bool autotune = false;
  std::map<std::string, int> tuned;
  TuneRecord last_tune;
  // hipGraph replay of engine calls (lc_engine_graph): the switch, the log of the call being
  // recorded, the graphs by call and what the last engine call did (0 eager, 1 recorded,
  // 2 replayed)
  bool graphs = false;
  bool tuning = false;  // the autotuner's candidate runs are never recorded
  LaunchLog* log = nullptr;
  std::map<std::string, GraphEntry> graph_cache;
  int last_graph = 0;
};

// Runs one kernel launch on the library stream; while an engine call is being recorded the launch
// is also kept for the graph.
void launch(const std::function<hipError_t(hipStream_t)>& f, const char* what);

// Runs one engine call (`call`, which ends with a stream synchronisation). With graphs on and no
// profiling, the first call of `key` (the engine, its arguments and the tile parameters) runs
// eagerly while its launches are logged and then captured into a graph; a later call with the same
// key replays the graph, unless the shared scratch or partial buffers were reallocated since, in
// which case it is recorded again.
void run_engine_call(const std::string& key, const std::function<void()>& call);

// Destroys every recorded graph.
void drop_graphs();

// Tiles-per-workgroup candidates of the autotuner, in timing order.
constexpr int kTileCandidates[3] = {4, 2, 1};

// Runs one engine call under the tile parameters for `key` (the engine and its shape). A
// tunable (mixed-form) call with the autotuner on, no override and no choice cached for `key`
// times every candidate on the call itself, twice each: an engine call overwrites all of its
// outputs, so repeating it gives the same results. The fastest is cached for `key` and its
// results are the ones left in the outputs.
void run_with_tiles(const std::string& key, bool tunable, const std::function<void()>& call);

Context& ctx();
void require_init();
// Default bound on an engine's scratch: 60% of the device memory free now (larger momentum
// chunks share each perambulator block between more momenta), at least 1 GiB.
long long default_scratch_budget();
// Message of the last failed C-ABI call (shared by every entry point).
std::string& last_error();

// A closing trace fused into a GEMM: out (+)= alpha * Tr[C * Y] for the product C being formed.
struct FusedTrace {
  LcView Y;
  void* out;
  long long os[3];
  double ar = 1.0, ai = 0.0;
  bool accumulate = false;
  int out_double = -1;  // -1: the program's compute type
  bool conj = false;    // deliver conj(alpha * Tr[C * Y])
};

inline FusedTrace fused(const LcView& Y, void* out, long long os0, long long os1, long long os2,
                        double ar = 1.0, double ai = 0.0) {
  FusedTrace f;
  f.Y = Y;
  f.out = out;
  f.os[0] = os0; f.os[1] = os1; f.os[2] = os2;
  f.ar = ar;
  f.ai = ai;
  return f;
}

// Padded block side of the exact form's digit storage: Nv rounded up to 32, 64 or 128.
inline int oz_pad(int nv) { return nv <= 32 ? 32 : nv <= 64 ? 64 : 128; }

// Digit output of an exact-form product: digit planes (strides in complex digit elements, P^2 per
// block) and the per-D-row exponents ([4][P] ints per matrix, strides in ints). At P = 128 a
// workgroup strip is half a D-row: e then receives the half-row exponents ([4 P][2] ints per
// matrix, INT_MIN for an all-zero half) with each half under its own, and Program::fix_digits
// makes the D-row form.
struct DigitOut {
  void* d = nullptr;
  int* e = nullptr;
  bool cf = false;  // the closing-factor form (int32 per value) instead of digits
  long long ds[3] = {0, 0, 0};
  long long es[3] = {0, 0, 0};
};

// A sequence of levels; each level is one grouped GEMM launch or one grouped trace launch
// (GEMM levels with fused traces are followed by their fixed-order reduction).
// Groups inside a level are independent; levels run in order on the library stream.
class Program {
 public:
  // F64 / F32: FP64 or FP32 matrix cores on FP64 / FP32 data. MX: FP32 data, GEMMs on the BF16
  // matrix cores with the exact three-way split (gemm_mx.hip); traces accumulate in FP64.
  // OZ: digit operands on the int8 matrix cores with exact accumulation (gemm_oz.hip), FP64
  // outputs, traces and closing factors.
  enum Mode { F64, F32, MX, OZ };
  explicit Program(int nv, bool dbl) : nv_(nv), dbl_(dbl), mode_(dbl ? F64 : F32) {}
  Program(int nv, Mode m) : nv_(nv), dbl_(m == F64 || m == OZ), mode_(m) {}

  // C = alpha * A * B over the batch; C may be nullptr when only fused traces are wanted.
  // MX mode: A and B are split views (A without LC_TV, B with LC_TV, so that both are read
  // along contiguous K) with phases in {+-1, +-i}; C is FP32 complex, and Cs (optional) receives
  // the same product in split storage, every block transposed when cs_trans.
  void gemm(const LcView& A, const LcView& B, void* C, long long cs0, long long cs1,
            long long cs2, int n0, int n1, int n2, double ar = 1.0, double ai = 0.0,
            const std::vector<FusedTrace>& traces = {}, void* Cs = nullptr, bool cs_trans = false);
  // Sandwich C = alpha * (L (x) 1) M (R (x) 1) over the batch, MX mode with Nv = 64 only
  // (sandwich_supported): L an A-role elemental view, M a B-role D-space view, R a B-role
  // elemental view. Each block is formed in one pass with L M_ac kept on the chip. With R2 (a
  // second B-role elemental view) and tr2, the same pass also forms (L (x) 1) M (R2 (x) 1) from the
  // same L M_ac and delivers only its trace tr2 (the output of tr2 is not scaled by alpha).
  void sandwich(const LcView& L, const LcView& M, const LcView& R, void* C, long long cs0, long long cs1,
                long long cs2, int n0, int n1, int n2, double ar = 1.0, double ai = 0.0,
                const std::vector<FusedTrace>& traces = {}, void* Cs = nullptr, bool cs_trans = false,
                const LcView* R2 = nullptr, const FusedTrace* tr2 = nullptr);
  bool sandwich_supported() const { return mode_ == MX && nv_ == 64; }
  // Workgroup order over a batch: by default axis 0 runs fastest. With last_axis_fastest, axis 2
  // does (e.g. the momenta of one (sink, source) pair next to each other, so that the shared
  // perambulator block is fetched from HBM once and from L2 for the other momenta).
  void set_last_axis_fastest(bool on) { last_axis_fastest_ = on; }
  // With last_axis_fastest: which of axes 0 and 1 runs next (default 0). Axis 1 suits a product
  // whose operand depends on axes 2 and 0 only, which neighbouring items then share.
  void set_second_axis(int axis) { second_axis_ = axis; }
  // OZ mode: C = alpha * A * B with A, B digit views (A read by rows, B by columns: B carries
  // LC_TV), phases powers of i. Outputs: FP64 C (optional, strides cs in complex elements of nv^2
  // blocks), digits with per-D-row exponents (optional; needs 4 P <= 256), fused traces (FP64 Y).
  // With twin_B (an elemental factor, at P = 128 an elemental B; no C): the twin product A twin_B
  // over the same batch, formed per item right after A B from the A operand already on the chip,
  // delivering only twin_traces.
  void gemm_oz(const LcView& A, const LcView& B, void* C, long long cs0, long long cs1, long long cs2,
               const DigitOut* dout, int n0, int n1, int n2, double ar = 1.0, double ai = 0.0,
               const std::vector<FusedTrace>& traces = {}, const LcView* twin_B = nullptr,
               const std::vector<FusedTrace>* twin_traces = nullptr);
  // OZ mode, chained: (A1 B1) B2 with A1 an elemental view, B1 a D-space view and B2 an elemental
  // view (4 P <= 256), outputs and traces (and twin) of the second product as in gemm_oz; the
  // intermediate A1 B1 stays on the chip (it becomes the second product's A operand in the stage
  // buffers, with its digits and row exponents exactly as a digit output would have them).
  void gemm_oz_chain(const LcView& A1, const LcView& B1, const LcView& B2, const DigitOut* dout, int n0, int n1,
                     int n2, const std::vector<FusedTrace>& traces = {}, const LcView* twin_B = nullptr,
                     const std::vector<FusedTrace>* twin_traces = nullptr);
  // OZ mode, P = 128: nmat consecutive digit matrices y (closing-factor form with cf) written with
  // half-row exponents eh (a DigitOut of gemm_oz) to the D-row exponents e, rescaling the half
  // of each row with the smaller exponent (one unit of rounding), as a level of its own.
  void fix_digits(void* y, const int* eh, int* e, long long nmat, bool cf);
  void trace(const LcView& X, const LcView& Y, void* out, long long os0, long long os1,
             long long os2, int n0, int n1, int n2, double ar = 1.0, double ai = 0.0,
             bool accumulate = false);
  // Close the current level.
  void barrier();
  // Upload descriptors and enqueue every level on the stream.
  void run(hipStream_t s);

  double gemm_flops() const { return gemm_flops_; }
  double trace_flops() const { return trace_flops_; }

 private:
  struct Fix {
    void* y;
    const int* eh;
    int* e;
    long long nmat;
    bool cf;
  };
  struct Level {
    bool is_trace;
    bool closed;
    int variant;  // MX kernel: 1 full-tile, 0 general; OZ: kOzGemm, kOzGemmRun, kOzFix
    bool xtr = false;  // OZ: a group with more than LC_MAX_FUSED_TRACES traces or elemental closing factors
    bool chain = false;  // OZ: a chained group (LcOzGroup::pre)
    size_t first;
    int count;
    long long total_wg;
    std::vector<int> jobs;  // reduce jobs run after this level
  };
  struct Patch {
    size_t group;
    int slot;
    size_t part_off;  // bytes into the partials buffer
    bool oz = false;
    bool twin = false;  // (oz) the group is in the twin array
  };
  static constexpr int kOzGemm = 10, kOzFix = 11, kOzGemmRun = 12;
  std::vector<LcOzGroup> ozs_, oztw_;
  std::vector<Fix> fixes_;
  void oz_items(const LcView& A, const LcView& B, void* C, const long long* cs, const DigitOut* dout, int n0,
                int n1, int n2, double ar, double ai, const std::vector<FusedTrace>& traces,
                const LcView* twin_B, const std::vector<FusedTrace>* twin_traces,
                const LcView* pre_A = nullptr, const LcView* pre_B = nullptr);
  int nv_;
  bool dbl_;
  Mode mode_;
  std::vector<LcGemmGroup> gemms_;
  std::vector<LcTraceGroup> traces_;
  std::vector<LcReduceJob> jobs_;
  std::vector<size_t> job_part_off_;
  std::vector<Patch> patches_;
  size_t partial_bytes_ = 0;
  std::vector<Level> levels_;
  double gemm_flops_ = 0, trace_flops_ = 0;
  Level* open_level(bool is_trace, int variant = 0);
  void gemm_items(const LcView& A, const LcView& B, void* C, long long cs0, long long cs1, long long cs2,
                  int n0, int n1, int n2, double ar, double ai, const std::vector<FusedTrace>& traces,
                  void* Cs, bool cs_trans, const LcView* B2 = nullptr, const LcView* B3 = nullptr,
                  const FusedTrace* trx = nullptr);
  void batch(const LcView& A, const LcView& B, const LcView* B2, void* C, long long cs0, long long cs1,
             long long cs2, int n0, int n1, int n2, double ar, double ai,
             const std::vector<FusedTrace>& traces, void* Cs, bool cs_trans, const LcView* B3 = nullptr,
             const FusedTrace* trx = nullptr);
  void add_reduce(Level* l, const FusedTrace& f, int n0, int n1, int n2, long long per_item, int slot);
  bool last_axis_fastest_ = false;
  int second_axis_ = 0;
};

}  // namespace lc
