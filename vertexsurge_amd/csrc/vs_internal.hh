// Internal declarations shared by the host (g++) and device (hipcc) halves of
// libvertexsurge.so.  Nothing here crosses the C ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/vertexsurge.h"

namespace vs {

// Thread-local last-error text behind vs_last_error().
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// Every C entry point returning vs_status is a function-try-block closed by
// VS_CATCH(name): a host allocation that throws (std::vector, std::thread,
// new) becomes VS_ERR_NOMEM, anything else VS_ERR_INTERNAL, so no C++
// exception leaves the library.  Device and pool buffers held by the call
// are RAII-owned and released while unwinding.
#define VS_CATCH(fn)                                                   \
  catch (const std::bad_alloc &) {                                     \
    ::vs::set_error("%s: out of host memory", fn);                     \
    return VS_ERR_NOMEM;                                               \
  }                                                                    \
  catch (const std::exception &e) {                                    \
    ::vs::set_error("%s: internal error: %s", fn, e.what());           \
    return VS_ERR_INTERNAL;                                            \
  }                                                                    \
  catch (...) {                                                        \
    ::vs::set_error("%s: internal error", fn);                         \
    return VS_ERR_INTERNAL;                                            \
  }

// Counter-based 64-bit mixer (splitmix64 finaliser over seed + ctr * golden).
// Every synthetic draw in the library is mix64(seed, counter), so results do
// not depend on thread count or scheduling.
static inline uint64_t mix64(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// floor(hi32(d) * n / 2^32): a uniform index in [0, n) for n < 2^32.
static inline uint32_t draw_index(uint64_t d, uint32_t n) {
  return static_cast<uint32_t>(((d >> 32) * static_cast<uint64_t>(n)) >> 32);
}

// Host CSR owned by the library (vs_csr_t).
struct HostCSR {
  uint32_t V = 0;
  uint64_t E = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> adj;
};

// Reverse (transpose) CSR on the host; lists come out sorted ascending.
void reverse_csr(const uint64_t *off, const uint32_t *adj, uint32_t V,
                 uint64_t E, std::vector<uint64_t> &roff,
                 std::vector<uint32_t> &radj);
// Union of a CSR and its reverse, deduplicated (Direction::Both).
void union_csr(const uint64_t *off, const uint32_t *adj, const uint64_t *roff,
               const uint32_t *radj, uint32_t V, std::vector<uint64_t> &uoff,
               std::vector<uint32_t> &uadj);

} // namespace vs
