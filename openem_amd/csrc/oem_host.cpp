// Host-side table helpers (include/openem_host.h).  Plain C++, no device code.

#include "../../include/openem_hip.h"
#include "../../include/openem_host.h"

#include <cstring>
#include <new>
#include <vector>

namespace {

inline uint64_t mix(uint64_t h, uint32_t w) {
    h ^= w;
    h *= 0x100000001b3ull;                  // FNV-1a prime, 64-bit
    return h ^ (h >> 29);
}

}  // namespace

extern "C" int oem_host_factorize(int64_t n, int ncols, const uint32_t *const *cols,
                                  const int32_t *stride, int32_t *inv, uint32_t *uniq,
                                  int64_t max_uniq, int64_t *n_uniq) {
    if (n < 0 || ncols <= 0 || ncols > 64 || !cols || !n_uniq || max_uniq <= 0 ||
        max_uniq > INT32_MAX || (n > 0 && (!inv || !uniq)))
        return OEM_EINVAL;
    int64_t st[64];
    for (int c = 0; c < ncols; ++c) {
        if (n > 0 && !cols[c]) return OEM_EINVAL;
        st[c] = stride ? stride[c] : 1;
        if (st[c] <= 0) return OEM_EINVAL;
    }
    *n_uniq = 0;
    if (n == 0) return OEM_OK;

    // open addressing over row numbers; load factor <= 1/2
    uint64_t cap = 16;
    const uint64_t want = (uint64_t)(max_uniq < n ? max_uniq : n) * 2;
    while (cap < want) cap <<= 1;
    std::vector<int32_t> slot;
    try {
        slot.assign(cap, -1);
    } catch (const std::bad_alloc &) {
        return OEM_ENOMEM;
    }
    const uint64_t mask = cap - 1;
    int64_t nu = 0;
    uint32_t row[64];
    int32_t prev = -1;                      // runs of equal rows are common (uniform regions)
    for (int64_t r = 0; r < n; ++r) {
        for (int c = 0; c < ncols; ++c) row[c] = cols[c][r * st[c]];
        if (prev >= 0 && std::memcmp(row, uniq + (size_t)prev * ncols, 4u * ncols) == 0) {
            inv[r] = prev;
            continue;
        }
        uint64_t h = 0xcbf29ce484222325ull;
        for (int c = 0; c < ncols; ++c) h = mix(h, row[c]);
        uint64_t s = h & mask;
        int32_t id;
        for (;;) {
            id = slot[s];
            if (id < 0) {
                if (nu >= max_uniq) {
                    *n_uniq = -1;
                    return OEM_OK;
                }
                id = (int32_t)nu++;
                std::memcpy(uniq + (size_t)id * ncols, row, 4u * ncols);
                slot[s] = id;
                break;
            }
            if (std::memcmp(row, uniq + (size_t)id * ncols, 4u * ncols) == 0) break;
            s = (s + 1) & mask;
        }
        inv[r] = id;
        prev = id;
    }
    *n_uniq = nu;
    return OEM_OK;
}
