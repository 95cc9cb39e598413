// Launch-size limit shared by the HIP translation units (no HIP dependency,
// so host-only tests compile it with a plain C++ compiler).
#pragma once

#include <cstdint>

namespace mafem {

// A dispatch counts its work-items in 32 bits (the grid size of the dispatch
// packet), so a launch of nblk workgroups of `threads` work-items each is
// representable only while nblk * threads <= 2^32 - 1.  Beyond that the
// grid wraps and the tail of the work is silently skipped.
inline bool launch_fits(uint64_t nblk, unsigned threads) {
  return threads > 0 && nblk <= 0xffffffffull / threads;
}

}  // namespace mafem
