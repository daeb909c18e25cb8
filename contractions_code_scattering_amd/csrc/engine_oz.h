// Exact form of the batched engines (engine_oz.cpp): entry points called by engine.cpp and
// engine_sigma.cpp when the mixed form selected by lc_set_mixed_form is the exact one.
#pragma once

namespace lc {

// rho + pi-pi set over a resident perambulator: back-to-back (d_phi [np][nt][nv][nv]) or explicit
// momentum sets (sets: d_phi [np][4][nt][nv][nv]); outputs as lc_engine_rho_pipi. Throws lc::Error.
void oz_engine_rho_pipi(bool sets, const void* d_tau, const void* d_phi, int nt, int nv, int np, int src0,
                        int src_step, int nsrc, long long max_scratch_bytes, void* d_raw, void* d_corr, void* d_avg);

// sigma / sigma-pi / rho-pi set (the inputs and raw slots of lc_engine_sigma_set): the bytes of the
// inputs in digits (ahead of the scratch area), the scratch of one momentum set, and the launches
// of the ten raw traces of every (tf, t0, p) into the context's stream (momentum sets in chunks of
// pc). Throws lc::Error.
size_t oz_sigma_head_bytes(int nt, int nv, int np);
size_t oz_sigma_bytes_per_momentum(int nt, int nv, int nsrc);
void oz_sigma_set(const void* d_tau, const void* d_el, int nt, int nv, int np, int idir, int src0, int src_step,
                  int nsrc, char* head, char* scratch, int pc, void* d_raw);

}  // namespace lc
