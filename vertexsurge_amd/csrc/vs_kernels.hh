// Hop kernels of the bit-matrix multi-source expansion, shared by the
// single-query engine (vs_vexpand.hip) and the batched engine (vs_batch.hip).
//
// One hop of the reference is the Boolean-semiring SpMM
//   out[si][v] |= base[si][u]  for every edge (u, v) and stack si
// (Expand<DenseVarStackMatrix>::expand_dense, src/objects/operator/
// expand.cpp:263-311), followed for Shortest by
//   o &= ~vis;  vis |= o;  visit_count = popcount(vis)
// (DenseVarStackMatrix::dedup_and_update_visit_and_count,
// src/objects/storage/bitmap.cpp:1054-1096), and the layers are unioned over
// [k_min, k_max] (VExpand::merge_to_MT, src/objects/operator/vexpand.cpp:298-354).
//
// Here a hop is either
//   pull: each destination vertex ORs the frontier words of its active
//         in-neighbours (gather; no atomics), with the dedup / visit / union
//         epilogue fused, and an early exit once every still-needed row bit
//         is found; or
//   push: each active frontier vertex scatters its words into a zeroed
//         accumulator N with 64-bit atomicOr, then an epilogue walks the
//         touched-vertex bitmap and applies the same dedup / visit / union.
// Both produce identical bits.  Frontier buffers and their active bitmaps are
// all-zero outside their active rows, and are cleared through the bitmap
// after use, so no hop pays for a full-matrix memset.
#pragma once

#include "vs_device.hh"

namespace vs {

// Per (hop, tile) counters, filled by the kernel that writes layer L_i.
struct HopCounters {
  ull pop;     // popcount of L_i
  ull nz;      // nonzero words of L_i
  ull act;     // vertices with a nonzero L_i row
  ull actdeg;  // sum of push-degree over those vertices
  ull te;      // sum over rows of push-degree * popcount(row): TE of hop i+1
  ull visnew;  // visited words that went from zero to nonzero
  ull rnew;    // result words that went from zero to nonzero
  ull items;   // push work items built from L_i (written by k_items)
};

struct HopArgs {
  uint32_t V;
  uint32_t S;      // total source rows of the query (valid-bit mask)
  uint32_t word0;  // global word index of this tile's word 0
  uint32_t pad;
  const uint64_t *poff;  // pull CSR: edges arriving at v
  const uint32_t *padj;
  const uint64_t *xoff;  // push CSR: edges leaving u
  const uint32_t *xadj;
  const uint64_t *F;     // frontier L_{i-1}, V*TW words
  const uint64_t *actbm; // its active-vertex bitmap, ceil(V/64) words
  uint64_t *Fn;          // L_i (all zero on entry)
  uint64_t *actbm_n;     // its bitmap (all zero on entry)
  uint64_t *N;           // push accumulator (all zero on entry and exit)
  uint64_t *touched;     // push touched bitmap (all zero on entry and exit)
  uint64_t *Vis;         // visited (Shortest) or nullptr (Any)
  uint64_t *R;           // result accumulator or nullptr
  const ull *items;      // push items built from actbm
  const ull *nitems;
  HopCounters *ctr;      // counters of L_i
};

struct LocalCtr {
  uint64_t pop = 0, nz = 0, act = 0, actdeg = 0, te = 0, visnew = 0, rnew = 0;
};

__device__ __forceinline__ void flush_ctr(LocalCtr &c, HopCounters *ctr) {
  const uint64_t pop = wave_sum(c.pop), nz = wave_sum(c.nz), act = wave_sum(c.act),
                 actdeg = wave_sum(c.actdeg), te = wave_sum(c.te),
                 visnew = wave_sum(c.visnew), rnew = wave_sum(c.rnew);
  if (lane_id() == 0) {
    if (pop) atomicAdd(&ctr->pop, static_cast<ull>(pop));
    if (nz) atomicAdd(&ctr->nz, static_cast<ull>(nz));
    if (act) atomicAdd(&ctr->act, static_cast<ull>(act));
    if (actdeg) atomicAdd(&ctr->actdeg, static_cast<ull>(actdeg));
    if (te) atomicAdd(&ctr->te, static_cast<ull>(te));
    if (visnew) atomicAdd(&ctr->visnew, static_cast<ull>(visnew));
    if (rnew) atomicAdd(&ctr->rnew, static_cast<ull>(rnew));
  }
}

// Writes one new word nb (nonzero) of L_i at index wi and updates Vis / R.
template <bool SHORTEST>
__device__ __forceinline__ void emit_word(const HopArgs &a, size_t wi, uint64_t nb,
                                          uint64_t vis, LocalCtr &c) {
  a.Fn[wi] = nb;
  if (SHORTEST) {
    a.Vis[wi] = vis | nb;
    c.visnew += (vis == 0);
  }
  if (a.R) {
    const uint64_t r = a.R[wi];
    if ((r | nb) != r) a.R[wi] = r | nb;
    c.rnew += (r == 0);
  }
  c.pop += __popcll(nb);
  c.nz += 1;
}

// ------------------------------------------------------------------- pull
// A wave owns 64 consecutive destination vertices (so it alone writes their
// active-bitmap word) and walks them one at a time.  Lanes = 64/TW groups x
// TW words; the groups split the in-neighbour list.  Neighbour ids are read
// 64 at a time, filtered by the frontier's active bitmap, compacted into LDS
// with ballot + mbcnt, and gathered with TW-word coalesced loads.
template <int TW, bool SHORTEST>
__global__ __launch_bounds__(kBlock) void k_pull(HopArgs a) {
  constexpr int G = 64 / TW;
  __shared__ uint32_t s_nbr[kWavesPerBlock][64];
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const int g = lane / TW, j = lane % TW;
  const uint64_t valid = valid_mask(a.word0 + j, a.S);
  const uint32_t nvb = (a.V + 63) / 64;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  const bool same_csr = a.xoff == a.poff;
  uint32_t *nbr = s_nbr[wid];
  LocalCtr c;
  for (uint32_t vb = blockIdx.x * kWavesPerBlock + wid; vb < nvb; vb += nwaves) {
    uint64_t actmask = 0;
    const uint32_t vend = min(64u, a.V - vb * 64);
    for (uint32_t i = 0; i < vend; ++i) {
      const uint32_t v = vb * 64 + i;
      const size_t wi = static_cast<size_t>(v) * TW + j;
      uint64_t vis = 0;
      if (SHORTEST) vis = a.Vis[wi];
      const uint64_t need = valid & ~vis;
      if (__ballot(need != 0) == 0) continue;
      const uint64_t b = a.poff[v], e = a.poff[v + 1];
      uint64_t acc = 0;
      for (uint64_t base = b; base < e; base += 64) {
        const uint64_t idx = base + lane;
        uint32_t u = 0;
        bool act = false;
        if (idx < e) {
          u = a.padj[idx];
          act = (a.actbm[u >> 6] >> (u & 63)) & 1;
        }
        const uint64_t m = __ballot(act);
        const uint32_t n = __popcll(m);
        if (act) nbr[lane_rank(m)] = u;
        wave_sync();
        uint32_t k = g;
        for (; k + 3 * G < n; k += 4 * G) {
          const uint32_t u0 = nbr[k], u1 = nbr[k + G], u2 = nbr[k + 2 * G],
                         u3 = nbr[k + 3 * G];
          const uint64_t x0 = a.F[static_cast<size_t>(u0) * TW + j];
          const uint64_t x1 = a.F[static_cast<size_t>(u1) * TW + j];
          const uint64_t x2 = a.F[static_cast<size_t>(u2) * TW + j];
          const uint64_t x3 = a.F[static_cast<size_t>(u3) * TW + j];
          acc |= (x0 | x1) | (x2 | x3);
        }
        for (; k < n; k += G) acc |= a.F[static_cast<size_t>(nbr[k]) * TW + j];
        wave_sync();
        acc = group_or<TW>(acc);
        if (__ballot((acc & need) != need) == 0) break;  // nothing left to find
      }
      acc = group_or<TW>(acc);
      const uint64_t nb = acc & need;
      if (g == 0 && nb) emit_word<SHORTEST>(a, wi, nb, vis, c);
      if (__ballot(g == 0 && nb != 0)) {
        actmask |= 1ull << i;
        const uint64_t deg = same_csr ? e - b : a.xoff[v + 1] - a.xoff[v];
        if (lane == 0) {
          c.act += 1;
          c.actdeg += deg;
        }
        if (g == 0) c.te += deg * __popcll(nb);
      }
    }
    if (lane == 0) a.actbm_n[vb] = actmask;
  }
  flush_ctr(c, a.ctr);
}

// ------------------------------------------------------------------- push
// One wave per work item (vertex, chunk of kPushChunk out-neighbours).
template <int TW>
__global__ __launch_bounds__(kBlock) void k_push(HopArgs a) {
  constexpr int G = 64 / TW;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const int g = lane / TW, j = lane % TW;
  const uint64_t nitems = *a.nitems;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t it = blockIdx.x * kWavesPerBlock + wid; it < nitems; it += nwaves) {
    const ull item = a.items[it];
    const uint32_t u = static_cast<uint32_t>(item);
    const uint32_t ch = static_cast<uint32_t>(item >> 32);
    const uint64_t ue = a.xoff[u + 1];
    const uint64_t b = a.xoff[u] + static_cast<uint64_t>(ch) * kPushChunk;
    const uint64_t e = min(b + kPushChunk, ue);
    const uint64_t f = a.F[static_cast<size_t>(u) * TW + j];
    for (uint64_t base = b; base < e; base += 64) {
      const uint64_t idx = base + lane;
      const uint32_t vv = idx < e ? a.xadj[idx] : 0u;
      const uint32_t n = static_cast<uint32_t>(min<uint64_t>(64, e - base));
      for (uint32_t k0 = 0; k0 < n; k0 += G) {
        const uint32_t k = k0 + g;
        const uint32_t v = __shfl(vv, k < n ? static_cast<int>(k) : 0);
        if (k < n) {
          if (f) atomicOr(reinterpret_cast<ull *>(&a.N[static_cast<size_t>(v) * TW + j]),
                          static_cast<ull>(f));
          if (j == 0)
            atomicOr(reinterpret_cast<ull *>(&a.touched[v >> 6]), 1ull << (v & 63));
        }
      }
    }
  }
}

// Push epilogue: for every touched vertex, consume N (re-zeroing it), apply
// dedup / visit / union, write L_i and its active bitmap word.
template <int TW, bool SHORTEST>
__global__ __launch_bounds__(kBlock) void k_push_epi(HopArgs a) {
  constexpr int G = 64 / TW;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const int g = lane / TW, j = lane % TW;
  const uint64_t valid = valid_mask(a.word0 + j, a.S);
  const uint32_t nvb = (a.V + 63) / 64;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  LocalCtr c;
  for (uint32_t vb = blockIdx.x * kWavesPerBlock + wid; vb < nvb; vb += nwaves) {
    uint64_t m = a.touched[vb];
    if (!m) continue;
    uint64_t actmask = 0;
    while (m) {
      const int i = nth_set_bit(m, g);
      m = drop_low_bits(m, G);
      uint64_t nb = 0;
      uint32_t v = 0;
      if (i >= 0) {
        v = vb * 64 + i;
        const size_t wi = static_cast<size_t>(v) * TW + j;
        const uint64_t n = a.N[wi];
        if (n) a.N[wi] = 0;
        uint64_t vis = 0;
        if (SHORTEST) vis = a.Vis[wi];
        nb = n & valid & ~vis;
        if (nb) emit_word<SHORTEST>(a, wi, nb, vis, c);
      }
      const uint64_t gb = group_bits<TW>(__ballot(nb != 0), g);
      if (i >= 0 && gb) {
        const uint64_t deg = a.xoff[v + 1] - a.xoff[v];
        if (j == 0) {
          actmask |= 1ull << i;
          c.act += 1;
          c.actdeg += deg;
        }
        c.te += deg * __popcll(nb);
      }
    }
    actmask = wave_or(actmask);
    if (lane == 0) {
      a.actbm_n[vb] = actmask;
      a.touched[vb] = 0;
    }
  }
  flush_ctr(c, a.ctr);
}

// Zero the active rows of a frontier buffer and its bitmap.
template <int TW>
__global__ __launch_bounds__(kBlock) void k_clear(uint64_t *F, uint64_t *actbm,
                                                 uint32_t V) {
  constexpr int G = 64 / TW;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const int g = lane / TW, j = lane % TW;
  const uint32_t nvb = (V + 63) / 64;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  for (uint32_t vb = blockIdx.x * kWavesPerBlock + wid; vb < nvb; vb += nwaves) {
    uint64_t m = actbm[vb];
    if (!m) continue;
    while (m) {
      const int i = nth_set_bit(m, g);
      m = drop_low_bits(m, G);
      if (i >= 0) F[(static_cast<size_t>(vb) * 64 + i) * TW + j] = 0;
    }
    if (lane == 0) actbm[vb] = 0;
  }
}

// Active bitmap -> push work items (vertex | chunk << 32); one thread per
// bitmap word, block-level exclusive scan, one atomicAdd per block.
__global__ __launch_bounds__(kBlock) void k_items(const uint64_t *actbm, uint32_t V,
                                                 const uint64_t *xoff, ull *items,
                                                 ull *nitems);

} // namespace vs
