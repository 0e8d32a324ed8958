// Push/pull scheduling: the plain density heuristic and the learned MLP.
//
// The reference's only adaptive rule is Expand<M>::run_internal's density
// switch p = visits / (V * S): CSR -> Dense below 0.4, Dense -> CSR above 0.6
// (src/objects/operator/expand.cpp:346-396).  Here the choice is per hop and
// per source tile between sparse scatter (push) and dense gather (pull); the
// heuristic compares the frontier's out-edge count with E / alpha, and the
// MLP predicts the time of each option from the same counters.
#pragma once

#include "vs_device.hh"

namespace vs {

struct SchedFeatures {
  uint32_t hop = 0;
  uint64_t rows = 0;   // source rows in the tile
  uint32_t tw = 0;     // tile width (words)
  uint32_t V = 0;
  uint64_t E = 0;
  uint64_t pop = 0;    // frontier popcount
  uint64_t act = 0;    // frontier vertices
  uint64_t actdeg = 0; // frontier out-edges (sum of degrees of active vertices)
  double vis_frac = 0; // fraction of visited words already nonzero
};

// Option indices of the MLP output (predicted log2 milliseconds): push and
// pull for one (hop, tile) decision; the whole query at tile width 16, 32,
// 64 for the query-level decision (feature rows with hop 0).
constexpr int kOptPush = 0;
constexpr int kOptPull = 1;
constexpr int kOptTW16 = 2;
constexpr int kOptTW32 = 3;
constexpr int kOptTW64 = 4;
constexpr int kNumOpts = 5;

// Query-level features of the tile-width decision.
struct QueryFeatures {
  uint64_t S = 0;       // source rows
  uint64_t degsum = 0;  // out-degree sum of the sources
  uint64_t vol2 = 0;    // 2-hop volume sum_s sum_{u in N(s)} deg(u)
  uint32_t k_max = 0;
  uint32_t V = 0;
  uint64_t E = 0;
};

inline int heuristic_decide(const SchedFeatures &f, const vs_tuning &t) {
  const uint64_t alpha = t.push_alpha ? t.push_alpha : 16;
  return (f.actdeg * alpha < f.E) ? 1 : 2;
}

// Feature vector (VS_MLP_IN floats) of one (hop, tile) decision.
void sched_features(const SchedFeatures &f, float *x);
// Feature vector of a query-level decision (x[3], the hop, is 0).
void query_features(const QueryFeatures &f, float *x);
// Tile width (16, 32 or 64, at most tw_max) with the lowest predicted time,
// or `fallback` when the MLP is unusable or does not predict it at least
// 0.1 log2 units (~7%) faster than the fallback.
uint32_t sched_tile_words(vs_sched_s *s, const QueryFeatures &f, uint32_t tw_max,
                          uint32_t fallback);
// 1 = push, 2 = pull; falls back to the heuristic if the MLP is unusable.
int sched_decide(vs_sched_s *s, const SchedFeatures &f, const vs_tuning &t);
// Device pointer of the scheduler's parameters (for in-graph inference).
const float *sched_params(vs_sched_s *s);

// MLP parameter layout (floats): W1[H][IN], b1[H], W2[H][H], b2[H],
// W3[OUT][H], b3[OUT].
constexpr int kIn = VS_MLP_IN, kHid = VS_MLP_HIDDEN, kOut = VS_MLP_OUT;
constexpr size_t kOffW1 = 0;
constexpr size_t kOffB1 = kOffW1 + static_cast<size_t>(kHid) * kIn;
constexpr size_t kOffW2 = kOffB1 + kHid;
constexpr size_t kOffB2 = kOffW2 + static_cast<size_t>(kHid) * kHid;
constexpr size_t kOffW3 = kOffB2 + kHid;
constexpr size_t kOffB3 = kOffW3 + static_cast<size_t>(kOut) * kHid;
constexpr size_t kNumParams = kOffB3 + kOut;
// After the parameters, on the device: the decision domain -- its row
// count n (as a float), the radius, then n feature rows of kIn floats.
constexpr uint32_t kMaxDomain = VS_SCHED_MAX_DOMAIN;
constexpr size_t kOffDomain = kNumParams;
constexpr size_t kDeviceFloats = kNumParams + 2 + static_cast<size_t>(kMaxDomain) * VS_MLP_IN;
static_assert(kHid == 64, "one wave lane per hidden unit");

// Whether feature vector x (read uniformly) lies within the domain stored
// after the parameters (always true without a domain).  One wave: lane k
// checks rows k, k + 64, ...
__device__ __forceinline__ bool mlp_in_domain_wave(const float *__restrict__ P,
                                                   const float *x) {
  const uint32_t n = static_cast<uint32_t>(P[kOffDomain]);
  if (n == 0) return true;
  const float r2 = P[kOffDomain + 1] * P[kOffDomain + 1];
  bool hit = false;
  for (uint32_t i = lane_id(); i < n; i += 64) {
    const float *row = P + kOffDomain + 2 + static_cast<size_t>(i) * kIn;
    float d = 0.f;
#pragma unroll
    for (int k = 0; k < kIn; ++k) {
      const float t = row[k] - x[k];
      d += t * t;
    }
    hit |= d <= r2;
  }
  return __ballot(hit) != 0;
}

// Forward pass of one sample by one wave: lane k is hidden unit k.  x is
// read uniformly; h1/h2 are the lane's activations, y (lanes < kOut) the
// output.  lds holds 2*kHid floats private to the wave.
__device__ __forceinline__ void mlp_forward_wave(const float *__restrict__ P,
                                                 const float *x, float *lds,
                                                 float &h1, float &h2, float &y) {
  const int k = lane_id();
  float a1 = P[kOffB1 + k];
#pragma unroll
  for (int i = 0; i < kIn; ++i) a1 += P[kOffW1 + k * kIn + i] * x[i];
  h1 = a1 > 0.f ? a1 : 0.f;
  lds[k] = h1;
  wave_sync();
  float a2 = P[kOffB2 + k];
  for (int m = 0; m < kHid; ++m) a2 += P[kOffW2 + k * kHid + m] * lds[m];
  h2 = a2 > 0.f ? a2 : 0.f;
  lds[kHid + k] = h2;
  wave_sync();
  y = 0.f;
  if (k < kOut) {
    float a3 = P[kOffB3 + k];
    for (int m = 0; m < kHid; ++m) a3 += P[kOffW3 + k * kHid + m] * lds[kHid + m];
    y = a3;
  }
  wave_sync();
}

} // namespace vs
