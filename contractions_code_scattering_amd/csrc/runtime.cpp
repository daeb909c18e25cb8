// Host runtime: context, Program builder/launcher.
#include "runtime.h"

#include <cstdlib>
#include <cstring>

#include "spin.h"

namespace lc {

Context& ctx() {
  static Context c;
  return c;
}

std::string& last_error() {
  static std::string e;
  return e;
}

void run_with_tiles(const std::string& key, bool tunable, const std::function<void()>& call) {
  Context& c = ctx();
  struct Restore {
    Context& c;
    ~Restore() { c.active_bpw = 0; }
  } restore{c};
  TuneRecord rec;
  rec.bpw = c.tile_bpw;
  if (!tunable || !c.autotune || c.tile_bpw > 0) {
    c.active_bpw = c.tile_bpw;
    call();
    c.last_tune = rec;
    return;
  }
  auto hit = c.tuned.find(key);
  if (hit != c.tuned.end()) {
    rec.bpw = hit->second;
    rec.how = 1;
    c.active_bpw = rec.bpw;
    call();
    c.last_tune = rec;
    return;
  }
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  hip_check(hipEventCreate(&ev.e[0]), "hipEventCreate");
  hip_check(hipEventCreate(&ev.e[1]), "hipEventCreate");
  struct Tuning {
    Context& c;
    ~Tuning() { c.tuning = false; }
  } tuning{c};
  c.tuning = true;
  int best = 0;
  for (int i = 0; i < 3; i++) {
    double t_min = 0;
    for (int rep = 0; rep < 2; rep++) {
      c.active_bpw = kTileCandidates[i];
      hip_check(hipEventRecord(ev.e[0], c.stream), "hipEventRecord");
      call();
      hip_check(hipEventRecord(ev.e[1], c.stream), "hipEventRecord");
      hip_check(hipEventSynchronize(ev.e[1]), "hipEventSynchronize");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "hipEventElapsedTime");
      if (rep == 0 || ms < t_min) t_min = ms;
    }
    rec.ms[i] = t_min;
    if (rec.ms[i] < rec.ms[best]) best = i;
  }
  c.tuning = false;
  // the outputs hold the last candidate's results; run the fastest again unless it was that one
  if (best != 2) {
    c.active_bpw = kTileCandidates[best];
    call();
  }
  c.tuned[key] = kTileCandidates[best];
  rec.bpw = kTileCandidates[best];
  rec.how = 2;
  c.last_tune = rec;
}

void launch(const std::function<hipError_t(hipStream_t)>& f, const char* what) {
  Context& c = ctx();
  hip_check(f(c.stream), what);
  if (c.log) c.log->launches.push_back(f);
}

static void destroy_entry(GraphEntry& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  for (DevBuf& b : g.bufs) b.release();
  g.bufs.clear();
}

void drop_graphs() {
  Context& c = ctx();
  for (auto& kv : c.graph_cache) destroy_entry(kv.second);
  c.graph_cache.clear();
}

void run_engine_call(const std::string& key, const std::function<void()>& call) {
  Context& c = ctx();
  c.last_graph = 0;
  if (!c.graphs || c.profile || c.tuning) {
    call();
    return;
  }
  char tail[64];
  std::snprintf(tail, sizeof tail, " bpw=%d dthor=%d", c.active_bpw, c.dthor_dagger ? 1 : 0);
  const std::string k = key + tail;
  auto hit = c.graph_cache.find(k);
  if (hit != c.graph_cache.end()) {
    GraphEntry& g = hit->second;
    if (g.scratch_gen == c.scratch.gen && g.partials_gen == c.partials.gen) {
      hip_check(hipGraphLaunch(g.exec, c.stream), "hipGraphLaunch");
      hip_check(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
      c.last_graph = 2;
      return;
    }
    destroy_entry(g);
    c.graph_cache.erase(hit);
  }
  // record: run eagerly (the results of this call) while logging the launches
  GraphEntry g;
  LaunchLog log;
  c.log = &log;
  try {
    call();
  } catch (...) {
    c.log = nullptr;
    for (DevBuf& b : log.bufs) b.release();
    throw;
  }
  c.log = nullptr;
  g.bufs = std::move(log.bufs);
  g.scratch_gen = c.scratch.gen;
  g.partials_gen = c.partials.gen;
  // capture the logged launches (kernel launches only: no copies, fills or host waits)
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (auto& f : log.launches) {
      e = f(c.stream);
      if (e != hipSuccess) break;
    }
    const hipError_t e2 = hipStreamEndCapture(c.stream, &graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    destroy_entry(g);
    hip_check(e, "graph capture");
  }
  c.graph_cache[k] = std::move(g);
  c.last_graph = 1;
}

long long default_scratch_budget() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return 64LL << 30;
  }
  // the engine's own scratch buffer counts as free: it is reused or grown
  const long long avail = (long long)free_b + (long long)ctx().scratch.cap;
  const long long b = avail / 10 * 6;
  return b > (1LL << 30) ? b : (1LL << 30);
}

void require_init() {
  if (!ctx().on) fail(LC_ERR_NOT_INIT, "contraction library is not initialised (call lc_init)");
}

Program::Level* Program::open_level(bool is_trace, int variant) {
  if (levels_.empty() || levels_.back().closed || levels_.back().is_trace != is_trace ||
      levels_.back().variant != variant) {
    Level l;
    l.is_trace = is_trace;
    l.closed = false;
    l.variant = variant;
    l.first = is_trace ? traces_.size() : (variant == kOzGemm || variant == kOzGemmRun) ? ozs_.size() : gemms_.size();
    l.count = 0;
    l.total_wg = 0;
    levels_.push_back(l);
  }
  return &levels_.back();
}

void Program::barrier() {
  if (!levels_.empty()) levels_.back().closed = true;
}

void Program::gemm(const LcView& A, const LcView& B, void* C, long long cs0, long long cs1,
                   long long cs2, int n0, int n1, int n2, double ar, double ai,
                   const std::vector<FusedTrace>& traces, void* Cs, bool cs_trans) {
  batch(A, B, nullptr, C, cs0, cs1, cs2, n0, n1, n2, ar, ai, traces, Cs, cs_trans);
}

void Program::sandwich(const LcView& L, const LcView& M, const LcView& R, void* C, long long cs0,
                       long long cs1, long long cs2, int n0, int n1, int n2, double ar, double ai,
                       const std::vector<FusedTrace>& traces, void* Cs, bool cs_trans, const LcView* R2,
                       const FusedTrace* tr2) {
  if (!sandwich_supported()) fail(LC_ERR_ARG, "fused sandwich needs the mixed-precision mode and Nv = 64");
  if (L.kind != LC_KRON || M.kind != LC_DENSE || R.kind != LC_KRON || (R2 && R2->kind != LC_KRON))
    fail(LC_ERR_ARG, "sandwich operands must be elemental, D-space, elemental");
  if ((R2 == nullptr) != (tr2 == nullptr)) fail(LC_ERR_ARG, "second sandwich factor needs its trace");
  batch(L, M, &R, C, cs0, cs1, cs2, n0, n1, n2, ar, ai, traces, Cs, cs_trans, R2, tr2);
}

void Program::batch(const LcView& A, const LcView& B, const LcView* B2, void* C, long long cs0, long long cs1,
                    long long cs2, int n0, int n1, int n2, double ar, double ai,
                    const std::vector<FusedTrace>& traces, void* Cs, bool cs_trans, const LcView* B3,
                    const FusedTrace* trx) {
  if (!last_axis_fastest_ || n2 <= 1) {
    gemm_items(A, B, C, cs0, cs1, cs2, n0, n1, n2, ar, ai, traces, Cs, cs_trans, B2, B3, trx);
    return;
  }
  // the same batch with its axes relabelled (2, 0, 1), or (2, 1, 0) with second_axis 1:
  // workgroups then walk axis 2 fastest, then the second axis
  const int o1 = second_axis_ == 1 ? 1 : 0, o2 = 1 - o1;
  auto rot = [&](LcView v) {
    const long long s[3] = {v.s[0], v.s[1], v.s[2]};
    v.s[0] = s[2]; v.s[1] = s[o1]; v.s[2] = s[o2];
    return v;
  };
  auto rot3 = [&](const long long* x, long long* y) {
    const long long t[3] = {x[0], x[1], x[2]};
    y[0] = t[2]; y[1] = t[o1]; y[2] = t[o2];
  };
  std::vector<FusedTrace> tr = traces;
  for (FusedTrace& f : tr) {
    f.Y = rot(f.Y);
    rot3(f.os, f.os);
  }
  LcView r2, r3;
  if (B2) r2 = rot(*B2);
  if (B3) r3 = rot(*B3);
  FusedTrace tx;
  if (trx) {
    tx = *trx;
    tx.Y = rot(tx.Y);
    rot3(trx->os, tx.os);
  }
  const long long cs[3] = {cs0, cs1, cs2};
  const int nn[3] = {n0, n1, n2};
  gemm_items(rot(A), rot(B), C, cs[2], cs[o1], cs[o2], nn[2], nn[o1], nn[o2], ar, ai, tr, Cs, cs_trans,
             B2 ? &r2 : nullptr, B3 ? &r3 : nullptr, trx ? &tx : nullptr);
}

void Program::gemm_items(const LcView& A, const LcView& B, void* C, long long cs0, long long cs1,
                         long long cs2, int n0, int n1, int n2, double ar, double ai,
                         const std::vector<FusedTrace>& traces, void* Cs, bool cs_trans, const LcView* B2,
                         const LcView* B3, const FusedTrace* trx) {
  if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
  if (A.kind == LC_KRON && B.kind == LC_KRON) fail(LC_ERR_ARG, "elemental x elemental product");
  if ((int)traces.size() > LC_MAX_FUSED_TRACES) fail(LC_ERR_ARG, "too many fused traces");
  if (!C && !Cs && traces.empty()) fail(LC_ERR_ARG, "product with neither output nor trace");
  {
    bool plain = kron_plain(A) && kron_plain(B) && (!B2 || kron_plain(*B2)) && (!B3 || kron_plain(*B3)) &&
                 (!trx || kron_plain(trx->Y));
    for (const FusedTrace& f : traces) plain = plain && kron_plain(f.Y);
    if (!plain) fail(LC_ERR_ARG, "an elemental view with a spin matrix outside the exact kernel's closing factors");
  }
  if (Cs && mode_ != MX) fail(LC_ERR_ARG, "split output outside the mixed-precision mode");
  int variant = 0;
  if (mode_ == MX) {
    if ((A.flags & LC_TV) || !(B.flags & LC_TV))
      fail(LC_ERR_ARG, "mixed-precision operands must be read along contiguous K (A row-major, B transposed)");
    for (const LcView* r : {B2, B3})
      if (r && (!(r->flags & LC_TV) || r->kind != LC_KRON))
        fail(LC_ERR_ARG, "sandwich right factor must be an elemental read along contiguous K");
    for (const LcView* v : {&A, &B})
      for (int a = 0; a < 4; a++)
        for (const double* ph : {v->ph_r[a], v->ph_c[a]})
          if (!((ph[1] == 0 && (ph[0] == 1 || ph[0] == -1)) || (ph[0] == 0 && (ph[1] == 1 || ph[1] == -1))))
            fail(LC_ERR_ARG, "mixed-precision operand phase (%g, %g) is not a power of i", ph[0], ph[1]);
    // complete 64 x 64 tiles and 16-byte alignment of every operand and output (split storage:
    // 12 B per complex element, FP32 output 8 B): full-tile kernel
    bool full = nv_ % 64 == 0;
    for (const LcView* v : {&A, &B, B2 ? B2 : &B, B3 ? B3 : &B}) {
      full = full && (reinterpret_cast<uintptr_t>(v->base) & 15) == 0;
      for (int d = 0; d < 3; d++) full = full && v->s[d] % 4 == 0;
    }
    full = full && (reinterpret_cast<uintptr_t>(C) & 15) == 0 && (reinterpret_cast<uintptr_t>(Cs) & 15) == 0;
    full = full && cs0 % 4 == 0 && cs1 % 4 == 0 && cs2 % 4 == 0;
    std::vector<const FusedTrace*> all;
    for (const FusedTrace& f : traces) all.push_back(&f);
    if (trx) all.push_back(trx);
    for (const FusedTrace* f : all) {  // FP32 trace operands, read four complex values at a time
      full = full && (reinterpret_cast<uintptr_t>(f->Y.base) & 15) == 0;
      for (int d = 0; d < 3; d++) full = full && f->Y.s[d] % 2 == 0;
    }
    variant = full ? 1 : 0;
    if (B2) {
      if (!full || nv_ != 64) fail(LC_ERR_ARG, "fused sandwich needs Nv = 64 and 16-byte aligned operands");
      variant = 2;
    }
  }
  Level* l = open_level(false, variant);
  LcGemmGroup g;
  std::memset(&g, 0, sizeof g);
  g.A = A;
  g.B = B;
  if (B2) g.B2 = *B2;
  if (B3) {
    if (!B2) fail(LC_ERR_ARG, "second right factor outside a sandwich");
    g.B3 = *B3;
    g.Yx = trx->Y;
    g.trx_alpha[0] = trx->ar;
    g.trx_alpha[1] = trx->ai;
  }
  g.C = C;
  g.Cs = Cs;
  g.cs_trans = cs_trans ? 1 : 0;
  g.cs[0] = cs0; g.cs[1] = cs1; g.cs[2] = cs2;
  g.alpha[0] = ar; g.alpha[1] = ai;
  g.n[0] = n0; g.n[1] = n1; g.n[2] = n2;
  const int tm = mode_ == MX ? 64 : lc_gemm_tile(nv_);
  g.tiles = (nv_ + tm - 1) / tm;
  g.wg_begin = l->total_wg;
  const long long items = (long long)n0 * n1 * n2;
  const long long tt = (long long)g.tiles * g.tiles;
  // workgroups (and trace pieces) per item: one per output tile, or one per `bpw` tiles in MX mode
  long long per_item = 16 * tt;
  if (mode_ == MX) {
    // 64 x 64 output tiles per workgroup: the engine call's setting, else 4 when a spin block is
    // one tile (Nv <= 64) and 1 otherwise
    const int want = ctx().active_bpw;
    g.bpw = want > 0 ? want : (tt == 1 ? 4 : 1);
    per_item = 16 * tt / g.bpw;
    // Items along batch axis 0 run next to each other. When only A changes along that axis, B is
    // shared by neighbouring items and stays in L2, while each A block is read by this item
    // alone: a workgroup then takes one block column, so that the four workgroups of the item
    // read each A block row at about the same time (one fetch from memory instead of one per
    // output tile).
    g.col_tiles = (g.bpw > 1 && A.s[0] != 0 && B.s[0] == 0) ? 1 : 0;
  }
  l->total_wg += items * per_item;
  l->count++;
  g.ntr = (int)traces.size();
  for (int t = 0; t < g.ntr; t++) {
    const FusedTrace& f = traces[t];
    g.Y[t] = f.Y;
    g.tr_alpha[t][0] = f.ar;
    g.tr_alpha[t][1] = f.ai;
    add_reduce(l, f, n0, n1, n2, per_item, t);
  }
  if (trx) add_reduce(l, *trx, n0, n1, n2, per_item, LC_MAX_FUSED_TRACES);
  gemms_.push_back(g);
  const int nterm = B2 ? (B3 ? 3 : 2) : (A.kind == LC_KRON || B.kind == LC_KRON) ? 1 : 4;
  gemm_flops_ += (double)items * 16 * nterm * 8.0 * nv_ * (double)nv_ * nv_;
}

// Fixed-order reduction of the per-workgroup pieces of one fused trace of the group being added;
// slot LC_MAX_FUSED_TRACES is the trace of a sandwich's second product.
void Program::add_reduce(Level* l, const FusedTrace& f, int n0, int n1, int n2, long long per_item, int slot) {
  const long long items = (long long)n0 * n1 * n2;
  const size_t bytes = (size_t)items * per_item * 16;
  patches_.push_back(Patch{gemms_.size(), slot, partial_bytes_});
  LcReduceJob j;
  std::memset(&j, 0, sizeof j);
  j.out = f.out;
  for (int d = 0; d < 3; d++) j.os[d] = f.os[d];
  j.n[0] = n0; j.n[1] = n1; j.n[2] = n2;
  j.nk = (int)per_item;
  j.accumulate = f.accumulate ? 1 : 0;
  j.conj = f.conj ? 1 : 0;
  j.out_double = f.out_double < 0 ? (dbl_ ? 1 : 0) : f.out_double;
  l->jobs.push_back((int)jobs_.size());
  jobs_.push_back(j);
  job_part_off_.push_back(partial_bytes_);
  partial_bytes_ += (bytes + 255) & ~size_t(255);
}

void Program::gemm_oz(const LcView& A, const LcView& B, void* C, long long cs0, long long cs1, long long cs2,
                      const DigitOut* dout, int n0, int n1, int n2, double ar, double ai,
                      const std::vector<FusedTrace>& traces, const LcView* twin_B,
                      const std::vector<FusedTrace>* twin_traces) {
  if (mode_ != OZ) fail(LC_ERR_ARG, "exact-form product outside the exact mode");
  const long long cs[3] = {cs0, cs1, cs2};
  if (!last_axis_fastest_ || n2 <= 1) {
    oz_items(A, B, C, cs, dout, n0, n1, n2, ar, ai, traces, twin_B, twin_traces);
    return;
  }
  // the same batch with its axes relabelled as in batch(): axis 2 fastest, then the second axis
  const int o1 = second_axis_ == 1 ? 1 : 0, o2 = 1 - o1;
  auto rot3 = [&](const long long* x, long long* y) {
    const long long t[3] = {x[0], x[1], x[2]};
    y[0] = t[2]; y[1] = t[o1]; y[2] = t[o2];
  };
  auto rot = [&](LcView v) {
    rot3(v.s, v.s);
    rot3(v.es, v.es);
    return v;
  };
  std::vector<FusedTrace> tr = traces, tr2 = twin_traces ? *twin_traces : std::vector<FusedTrace>{};
  for (std::vector<FusedTrace>* v : {&tr, &tr2})
    for (FusedTrace& f : *v) {
      f.Y = rot(f.Y);
      rot3(f.os, f.os);
    }
  LcView tb;
  if (twin_B) tb = rot(*twin_B);
  long long rc[3];
  rot3(cs, rc);
  DigitOut d;
  if (dout) {
    d = *dout;
    rot3(dout->ds, d.ds);
    rot3(dout->es, d.es);
  }
  const int nn[3] = {n0, n1, n2};
  oz_items(rot(A), rot(B), C, rc, dout ? &d : nullptr, nn[2], nn[o1], nn[o2], ar, ai, tr, twin_B ? &tb : nullptr,
           twin_B ? &tr2 : nullptr);
}

void Program::gemm_oz_chain(const LcView& A1, const LcView& B1, const LcView& B2, const DigitOut* dout, int n0, int n1,
                            int n2, const std::vector<FusedTrace>& traces, const LcView* twin_B,
                            const std::vector<FusedTrace>* twin_traces) {
  if (mode_ != OZ) fail(LC_ERR_ARG, "exact-form product outside the exact mode");
  // the chained product's A: the first product, never stored (a placeholder view)
  const LcView chained = make_view(nullptr, 0, 0, 0, LC_DENSE, 0);
  const long long cs[3] = {0, 0, 0};
  if (!last_axis_fastest_ || n2 <= 1) {
    oz_items(chained, B2, nullptr, cs, dout, n0, n1, n2, 1.0, 0.0, traces, twin_B, twin_traces, &A1, &B1);
    return;
  }
  // the axes relabelled as in gemm_oz
  const int o1 = second_axis_ == 1 ? 1 : 0, o2 = 1 - o1;
  auto rot3 = [&](const long long* x, long long* y) {
    const long long t[3] = {x[0], x[1], x[2]};
    y[0] = t[2]; y[1] = t[o1]; y[2] = t[o2];
  };
  auto rot = [&](LcView v) {
    rot3(v.s, v.s);
    rot3(v.es, v.es);
    return v;
  };
  std::vector<FusedTrace> tr = traces, tr2 = twin_traces ? *twin_traces : std::vector<FusedTrace>{};
  for (std::vector<FusedTrace>* v : {&tr, &tr2})
    for (FusedTrace& f : *v) {
      f.Y = rot(f.Y);
      rot3(f.os, f.os);
    }
  LcView tb;
  if (twin_B) tb = rot(*twin_B);
  DigitOut d;
  if (dout) {
    d = *dout;
    rot3(dout->ds, d.ds);
    rot3(dout->es, d.es);
  }
  const LcView a1 = rot(A1), b1 = rot(B1);
  const int nn[3] = {n0, n1, n2};
  oz_items(chained, rot(B2), nullptr, cs, dout ? &d : nullptr, nn[2], nn[o1], nn[o2], 1.0, 0.0, tr,
           twin_B ? &tb : nullptr, twin_B ? &tr2 : nullptr, &a1, &b1);
}

void Program::oz_items(const LcView& A, const LcView& B, void* C, const long long* cs, const DigitOut* dout, int n0,
                       int n1, int n2, double ar, double ai, const std::vector<FusedTrace>& traces,
                       const LcView* twin_B, const std::vector<FusedTrace>* twin_traces, const LcView* pre_A,
                       const LcView* pre_B) {
  if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
  const int P = oz_pad(nv_);
  if (nv_ > 128) fail(LC_ERR_ARG, "exact form supports nv <= 128 (got %d)", nv_);
  if (A.kind == LC_KRON && B.kind == LC_KRON) fail(LC_ERR_ARG, "elemental x elemental product");
  if ((int)traces.size() > LC_OZ_MAX_TRACES) fail(LC_ERR_ARG, "too many fused traces");
  if (!C && !dout && traces.empty()) fail(LC_ERR_ARG, "product with neither output nor trace");
  if (!kron_plain(A) || !kron_plain(B) || (twin_B && !kron_plain(*twin_B)))
    fail(LC_ERR_ARG, "an elemental operand with a spin matrix");
  // a chained product (A = pre_A pre_B, formed on the chip): elemental pre_A times dense pre_B, then
  // this product dense x elemental, with the whole D-row in one workgroup strip
  const bool chain = pre_A != nullptr;
  if (chain) {
    if (!pre_B || pre_A->kind != LC_KRON || pre_B->kind != LC_DENSE || B.kind != LC_KRON || C || 4 * P > 256)
      fail(LC_ERR_ARG, "a chained product is (elemental x dense) x elemental with 4 P <= 256 and no FP64 output");
    if (!kron_plain(*pre_A) || (pre_A->flags & LC_TV) || !(pre_B->flags & LC_TV) || !pre_A->ex || !pre_B->ex)
      fail(LC_ERR_ARG, "a chained first product's operands must be read along contiguous K, with exponents");
    if (pre_B->ebs && !(pre_B->flags & LC_TS))
      fail(LC_ERR_ARG, "exact-form dense operand lines cross stored block rows");
    for (const LcView* v : {pre_A, pre_B})
      for (int a = 0; a < 4; a++)
        for (const double* ph : {v->ph_r[a], v->ph_c[a]})
          if (!((ph[1] == 0 && (ph[0] == 1 || ph[0] == -1)) || (ph[0] == 0 && (ph[1] == 1 || ph[1] == -1))))
            fail(LC_ERR_ARG, "exact-form operand phase (%g, %g) is not a power of i", ph[0], ph[1]);
  }
  if ((!chain && (A.flags & LC_TV)) || !(B.flags & LC_TV))
    fail(LC_ERR_ARG, "exact-form operands must be read along contiguous K (A by rows, B by columns)");
  if ((!chain && !A.ex) || !B.ex) fail(LC_ERR_ARG, "exact-form operand without exponents");
  for (const LcView* v : {&A, &B})
    for (int a = 0; a < 4; a++)
      for (const double* ph : {v->ph_r[a], v->ph_c[a]})
        if (!((ph[1] == 0 && (ph[0] == 1 || ph[0] == -1)) || (ph[0] == 0 && (ph[1] == 1 || ph[1] == -1))))
          fail(LC_ERR_ARG, "exact-form operand phase (%g, %g) is not a power of i", ph[0], ph[1]);
  const int kind = A.kind == LC_KRON ? 1 : B.kind == LC_KRON ? 2 : 0;
  // a dense K sum runs over the four blocks of a line: per-row exponents must then be those of
  // one stored block row (A without, B with the spin-block transpose)
  if (kind == 0 && ((A.ebs && (A.flags & LC_TS)) || (B.ebs && !(B.flags & LC_TS))))
    fail(LC_ERR_ARG, "exact-form dense operand lines cross stored block rows");
  // closing factors: all FP64, or all in the closing-factor form (int32 per value under line
  // exponents) read in the orientation of the product (LC_TV), dense or an elemental with a spin
  // matrix (one exponent per line index)
  auto check_traces = [](const std::vector<FusedTrace>& trs) {
    for (const FusedTrace& f : trs) {
      if ((f.Y.ex != nullptr) != (trs[0].Y.ex != nullptr))
        fail(LC_ERR_ARG, "exact-form closing factors of one product mix FP64 and the closing-factor form");
      if (f.Y.ex && !(f.Y.flags & LC_TV))
        fail(LC_ERR_ARG, "a closing factor in the closing-factor form must be read along its stored rows (LC_TV)");
      if (f.Y.kind == LC_KRON && (f.Y.ebs || (!f.Y.ex && !kron_plain(f.Y))))
        fail(LC_ERR_ARG, "an elemental closing factor with per-block exponents or an FP64 one with a spin matrix");
    }
  };
  check_traces(traces);
  const int NW = P == 32 ? 4 : 8;
  const int nseg = (4 * P) / (NW * 32);
  if (dout && (!dout->d || !dout->e)) fail(LC_ERR_ARG, "exact-form digit output without planes or exponents");
  // a twin reuses the A operand the kernel staged for this product: every K step of it must still
  // be in the stage buffers (an elemental factor, one block of K <= 64; at P = 128 an elemental B,
  // whose 4 steps of 2 chunks fill them), and nothing of the product may use the FP64 tile that
  // overlaps them (no FP64 output, closing factors in the closing-factor form)
  if (twin_B) {
    if (!twin_traces || twin_traces->empty() || (int)twin_traces->size() > LC_MAX_FUSED_TRACES)
      fail(LC_ERR_ARG, "a twin product needs 1..%d fused traces", LC_MAX_FUSED_TRACES);  // (the kernel's twin path)
    if (kind == 0 || (P > 64 && kind != 2))
      fail(LC_ERR_ARG, "twin products need an elemental factor (at P = 128, an elemental B)");
    if ((twin_B->kind == LC_KRON) != (B.kind == LC_KRON) || !(twin_B->flags & LC_TV) || !twin_B->ex)
      fail(LC_ERR_ARG, "a twin's B must be read like B (same kind, by columns, with exponents)");
    for (int a = 0; a < 4; a++)
      for (const double* ph : {twin_B->ph_r[a], twin_B->ph_c[a]})
        if (!((ph[1] == 0 && (ph[0] == 1 || ph[0] == -1)) || (ph[0] == 0 && (ph[1] == 1 || ph[1] == -1))))
          fail(LC_ERR_ARG, "exact-form operand phase (%g, %g) is not a power of i", ph[0], ph[1]);
    if (C) fail(LC_ERR_ARG, "a product with a twin cannot have an FP64 output");
    if (!traces.empty() && !traces[0].Y.ex)
      fail(LC_ERR_ARG, "a product with a twin needs its closing factors in the closing-factor form");
    check_traces(*twin_traces);
    if (!(*twin_traces)[0].Y.ex) fail(LC_ERR_ARG, "a twin's closing factors must be in the closing-factor form");
  }
  // runs of items along axis 0 that share A: a dense x dense product (every K step kept on the
  // chip, 4 P / 32 <= 8 stage buffers of 16 KB), A independent of axis 0, nothing that uses the
  // FP64 tile (it overlaps the stage buffers)
  const bool run_ok = kind == 0 && n0 > 1 && P <= 64 && A.s[0] == 0 && A.es[0] == 0 && !C && !twin_B &&
                      (traces.empty() || traces[0].Y.ex);
  // a launch holds runs only or single items only: the persistent workgroups split the launch's
  // logical workgroups evenly, which balances the work only when they are alike
  Level* l = open_level(false, run_ok ? kOzGemmRun : kOzGemm);
  // the kernel's extended trace path: more than LC_MAX_FUSED_TRACES traces, or elemental closing factors
  if ((int)traces.size() > LC_MAX_FUSED_TRACES) l->xtr = true;
  for (const std::vector<FusedTrace>* v : {&traces, twin_traces})
    if (v)
      for (const FusedTrace& f : *v)
        if (f.Y.kind == LC_KRON) l->xtr = true;
  LcOzGroup g;
  std::memset(&g, 0, sizeof g);
  g.A = A;
  g.B = B;
  g.C = C;
  for (int d = 0; d < 3; d++) g.cs[d] = cs[d];
  if (dout) {
    g.Cd = dout->d;
    g.Ce = dout->e;
    g.cf = dout->cf ? 1 : 0;
    for (int d = 0; d < 3; d++) {
      g.cds[d] = dout->ds[d];
      g.ces[d] = dout->es[d];
    }
  }
  g.alpha[0] = ar;
  g.alpha[1] = ai;
  g.n[0] = n0; g.n[1] = n1; g.n[2] = n2;
  g.kind = kind;
  g.nseg = nseg;
  g.twin = -1;
  g.pre = -1;
  g.arun = run_ok ? n0 : 1;
  g.wg_begin = l->total_wg;
  const long long items = (long long)n0 * n1 * n2;
  const long long per_item = (long long)(4 * P / 32) * nseg;
  l->total_wg += items / g.arun * per_item;
  l->count++;
  g.ntr = (int)traces.size();
  for (int t = 0; t < g.ntr; t++) {
    const FusedTrace& f = traces[t];
    g.Y[t] = f.Y;
    g.tr_alpha[t][0] = f.ar;
    g.tr_alpha[t][1] = f.ai;
    add_reduce(l, f, n0, n1, n2, per_item, t);
    patches_.back().group = ozs_.size();
    patches_.back().oz = true;
  }
  const int nterm = kind == 0 ? 4 : 1;
  gemm_flops_ += (double)items * 16 * nterm * 8.0 * nv_ * (double)nv_ * nv_;
  if (twin_B) {
    LcOzGroup t = g;
    t.B = *twin_B;
    t.Cd = nullptr;
    t.Ce = nullptr;
    t.cf = 0;
    t.alpha[0] = 1.0;
    t.alpha[1] = 0.0;
    t.ntr = (int)twin_traces->size();
    for (int k = 0; k < t.ntr; k++) {
      const FusedTrace& f = (*twin_traces)[k];
      t.Y[k] = f.Y;
      t.tr_alpha[k][0] = f.ar;
      t.tr_alpha[k][1] = f.ai;
      add_reduce(l, f, n0, n1, n2, per_item, k);
      patches_.back().group = oztw_.size();
      patches_.back().oz = true;
      patches_.back().twin = true;
    }
    g.twin = (int)oztw_.size();
    oztw_.push_back(t);
    gemm_flops_ += (double)items * 16 * 8.0 * nv_ * (double)nv_ * nv_;
  }
  if (chain) {
    // the first product in the side array: elemental x dense, no outputs of its own
    LcOzGroup f;
    std::memset(&f, 0, sizeof f);
    f.A = *pre_A;
    f.B = *pre_B;
    f.alpha[0] = 1.0;
    f.n[0] = n0; f.n[1] = n1; f.n[2] = n2;
    f.kind = 1;
    f.nseg = 1;
    f.twin = -1;
    f.pre = -1;
    f.arun = 1;
    g.pre = (int)oztw_.size();
    oztw_.push_back(f);
    l->chain = true;
    gemm_flops_ += (double)items * 16 * 8.0 * nv_ * (double)nv_ * nv_;
  }
  ozs_.push_back(g);
}

void Program::fix_digits(void* y, const int* eh, int* e, long long nmat, bool cf) {
  if (mode_ != OZ || oz_pad(nv_) != 128) fail(LC_ERR_ARG, "half-row digit fix-up outside the exact mode at P = 128");
  if (nmat <= 0) return;
  Level* l = open_level(false, kOzFix);
  l->first = fixes_.size();
  l->count = 1;
  fixes_.push_back(Fix{y, eh, e, nmat, cf});
  l->closed = true;
}

void Program::trace(const LcView& X, const LcView& Y, void* out, long long os0, long long os1,
                    long long os2, int n0, int n1, int n2, double ar, double ai, bool accumulate) {
  if (n0 <= 0 || n1 <= 0 || n2 <= 0) return;
  if (!kron_plain(X) || !kron_plain(Y)) fail(LC_ERR_ARG, "an elemental view with a spin matrix in a trace");
  Level* l = open_level(true);
  LcTraceGroup g;
  std::memset(&g, 0, sizeof g);
  g.X = X;
  g.Y = Y;
  g.out = out;
  g.os[0] = os0; g.os[1] = os1; g.os[2] = os2;
  g.alpha[0] = ar; g.alpha[1] = ai;
  g.n[0] = n0; g.n[1] = n1; g.n[2] = n2;
  g.accumulate = accumulate ? 1 : 0;
  g.wg_begin = l->total_wg;
  const long long items = (long long)n0 * n1 * n2;
  l->total_wg += items;
  l->count++;
  traces_.push_back(g);
  const int npair = (X.kind == LC_KRON || Y.kind == LC_KRON) ? 4 : 16;
  trace_flops_ += (double)items * npair * 8.0 * nv_ * nv_;
}

void Program::run(hipStream_t s) {
  Context& c = ctx();
  // The descriptor and partial buffers are reused by the next program: wait for the previous one.
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (partial_bytes_) c.partials.ensure(partial_bytes_);
  char* pbase = static_cast<char*>(c.partials.p);
  for (const Patch& p : patches_) {
    double* q = reinterpret_cast<double*>(pbase + p.part_off);
    if (p.oz) (p.twin ? oztw_ : ozs_)[p.group].part[p.slot] = q;
    else if (p.slot < LC_MAX_FUSED_TRACES) gemms_[p.group].part[p.slot] = q;
    else gemms_[p.group].partx = q;
  }
  for (size_t j = 0; j < jobs_.size(); j++) jobs_[j].part = reinterpret_cast<const double*>(pbase + job_part_off_[j]);
  // reduce jobs of each level, contiguous, with their thread offsets
  std::vector<LcReduceJob> jobs;
  std::vector<long long> begins;
  std::vector<std::pair<size_t, long long>> level_jobs;  // (first job, total threads) per level
  for (Level& l : levels_) {
    long long tot = 0;
    const size_t first = jobs.size();
    for (int j : l.jobs) {
      begins.push_back(tot);
      jobs.push_back(jobs_[j]);
      tot += (long long)jobs_[j].n[0] * jobs_[j].n[1] * jobs_[j].n[2];
    }
    level_jobs.push_back({first, tot});
  }
  const size_t gb = gemms_.size() * sizeof(LcGemmGroup);
  const size_t tb = traces_.size() * sizeof(LcTraceGroup);
  const size_t jb = jobs.size() * sizeof(LcReduceJob);
  const size_t bb = begins.size() * sizeof(long long);
  const size_t ob = ozs_.size() * sizeof(LcOzGroup);
  const size_t wb = oztw_.size() * sizeof(LcOzGroup);
  const size_t toff = (gb + 255) & ~size_t(255);
  const size_t joff = (toff + tb + 255) & ~size_t(255);
  const size_t boff = (joff + jb + 255) & ~size_t(255);
  const size_t ooff = (boff + bb + 255) & ~size_t(255);
  // descriptors: the shared buffer, or, while an engine call is recorded for a graph, a buffer of
  // this program's own that lives as long as the graph
  DevBuf* dbuf = &c.groups;
  if (c.log) {
    c.log->bufs.emplace_back();
    dbuf = &c.log->bufs.back();
  }
  const size_t woff = (ooff + ob + 255) & ~size_t(255);
  dbuf->ensure(woff + wb + 256);
  char* base = static_cast<char*>(dbuf->p);
  if (gb) hip_check(hipMemcpyAsync(base, gemms_.data(), gb, hipMemcpyHostToDevice, s), "upload gemm groups");
  if (tb) hip_check(hipMemcpyAsync(base + toff, traces_.data(), tb, hipMemcpyHostToDevice, s), "upload trace groups");
  if (jb) hip_check(hipMemcpyAsync(base + joff, jobs.data(), jb, hipMemcpyHostToDevice, s), "upload reduce jobs");
  if (bb) hip_check(hipMemcpyAsync(base + boff, begins.data(), bb, hipMemcpyHostToDevice, s), "upload reduce offsets");
  if (ob) hip_check(hipMemcpyAsync(base + ooff, ozs_.data(), ob, hipMemcpyHostToDevice, s), "upload exact-form groups");
  if (wb) hip_check(hipMemcpyAsync(base + woff, oztw_.data(), wb, hipMemcpyHostToDevice, s), "upload exact-form twins");
  const LcOzGroup* dozs = reinterpret_cast<const LcOzGroup*>(base + ooff);
  const LcOzGroup* dtw = reinterpret_cast<const LcOzGroup*>(base + woff);
  const LcReduceJob* dj = reinterpret_cast<const LcReduceJob*>(base + joff);
  const long long* db = reinterpret_cast<const long long*>(base + boff);
  // host vectors may be pageable: make sure the copies have landed before they can change
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  const LcGemmGroup* dg = reinterpret_cast<const LcGemmGroup*>(base);
  const LcTraceGroup* dt = reinterpret_cast<const LcTraceGroup*>(base + toff);
  const bool prof = c.profile;
  size_t nev = 0;
  if (prof) {
    while (c.ev_pool.size() < 2 * levels_.size()) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      c.ev_pool.push_back(e);
    }
  }
  for (size_t li = 0; li < levels_.size(); li++) {
    const Level& l = levels_[li];
    if (l.count == 0) continue;
    if (l.total_wg > 0x7ffffff0LL) fail(LC_ERR_ARG, "launch too large (%lld workgroups)", l.total_wg);
    if (prof) hip_check(hipEventRecord(c.ev_pool[nev], s), "hipEventRecord");
    const bool dbl = dbl_, is_trace = l.is_trace, mx = mode_ == MX;
    const int variant = l.variant, count = l.count, nv = nv_;
    const long long total_wg = l.total_wg;
    if (variant == kOzFix) {
      const Fix f = fixes_[l.first];
      launch([=](hipStream_t st) { return lc_launch_oz_fix(f.y, f.eh, f.e, f.nmat, 128, f.cf ? 1 : 0, st); },
             "digit fix-up");
    } else if (variant == kOzGemm || variant == kOzGemmRun) {
      const LcOzGroup* og = dozs + l.first;
      const int P = oz_pad(nv);
      const int run = variant == kOzGemmRun ? 1 : 0, xtr = l.xtr ? 1 : 0, chain = l.chain ? 1 : 0;
      if (chain && xtr) fail(LC_ERR_ARG, "a launch with chained groups and the extended trace path");
      launch([=](hipStream_t st) { return lc_launch_gemm_oz(og, count, dtw, total_wg, nv, P, run, xtr, chain, st); },
             "exact gemm launch");
    } else {
      const LcTraceGroup* tg = dt + l.first;
      const LcGemmGroup* gg = dg + l.first;
      launch(
          [=](hipStream_t st) {
            return is_trace ? lc_launch_trace(dbl, tg, count, total_wg, nv, st)
                            : (mx ? lc_launch_gemm_mx(variant, gg, count, total_wg, nv, st)
                                  : lc_launch_gemm(dbl, gg, count, total_wg, nv, st));
          },
          l.is_trace ? "trace launch" : "gemm launch");
    }
    if (prof) hip_check(hipEventRecord(c.ev_pool[nev + 1], s), "hipEventRecord");
    nev += 2;
    if (!l.jobs.empty()) {
      const LcReduceJob* rj = dj + level_jobs[li].first;
      const long long* rb = db + level_jobs[li].first;
      const int nj = (int)l.jobs.size();
      const long long nthr = level_jobs[li].second;
      launch([=](hipStream_t st) { return lc_launch_reduce(rj, rb, nj, nthr, st); }, "reduce launch");
    }
  }
  if (prof) {
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    size_t k = 0;
    for (const Level& l : levels_) {
      if (l.count == 0) continue;
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, c.ev_pool[k], c.ev_pool[k + 1]), "hipEventElapsedTime");
      if (l.is_trace) { c.times.trace_ms += ms; c.times.trace_launches++; }
      else { c.times.gemm_ms += ms; c.times.gemm_launches++; }
      k += 2;
    }
    c.times.gemm_flops += gemm_flops_;
    c.times.trace_flops += trace_flops_;
  }
}

}  // namespace lc
