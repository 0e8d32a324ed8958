// Batched VExpand (many queries packed into shared tiles) and the autotuner.
#include "vs_kernels.hh"

using namespace vs;

extern "C" {

vs_status vs_vexpand_batch(vs_graph_t g, const vs_query *queries, uint32_t nq,
                           const vs_batch_params *p, uint64_t *row_counts,
                           uint64_t *row_digests, vs_batch_stats *stats) {
  (void)g; (void)queries; (void)nq; (void)p; (void)row_counts; (void)row_digests;
  (void)stats;
  set_error("vs_vexpand_batch: not built yet");
  return VS_ERR_UNSUPPORTED;
}

vs_status vs_autotune_vexpand(vs_graph_t g, const uint32_t *sources, uint32_t S,
                              uint32_t k_max, int32_t path_type, vs_tuning *best,
                              double *best_ms) {
  (void)g; (void)sources; (void)S; (void)k_max; (void)path_type; (void)best; (void)best_ms;
  set_error("vs_autotune_vexpand: not built yet");
  return VS_ERR_UNSUPPORTED;
}

} // extern "C"
