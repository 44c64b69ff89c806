// Tandem-repeat length of the PCR indel model, shared by the host adapter and
// the device preprocessing kernel (ph/pairhmm_internal.cpp:260-314).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rvc {

constexpr int kMaxRepeatUnit = 8;   // longest repeat unit examined
constexpr int kMaxRepeatLen = 20;   // repeat counts are capped here

__host__ __device__ inline bool same_bytes(const uint8_t* a, const uint8_t* b, int n) {
  for (int x = 0; x < n; ++x)
    if (a[x] != b[x]) return false;
  return true;
}

// How many whole copies of unit = s[u .. u+ul) tile s[t .. t+tl): from the
// front (leading) or from the back, stopping at the first mismatch.
__host__ __device__ inline int count_copies(const uint8_t* s, int u, int ul, int t, int tl, bool leading) {
  if (tl == 0) return 0;
  int n = 0;
  const int last = tl - ul;
  if (leading) {
    for (int p = 0; p <= last; p += ul) {
      if (!same_bytes(s + t + p, s + u, ul)) break;
      ++n;
    }
  } else {
    for (int p = last; p >= 0; p -= ul) {
      if (!same_bytes(s + t + p, s + u, ul)) break;
      ++n;
    }
  }
  return n;
}

// Tandem-repeat length around base 'pos' of read s[0 .. n) as Rovaca computes
// it. The backward unit is recorded with the start pos - len - 1, as there;
// that start never equals the forward unit's start pos + 1, so the repeat
// length is always the forward unit's count on both sides.
__host__ __device__ inline int repeat_length(const uint8_t* s, int n, int pos) {
  int bw_count = 0, bw_start = pos, bw_len = 1;
  for (int ul = 1; ul <= kMaxRepeatUnit; ++ul) {
    if (pos + 1 - ul < 0) break;
    bw_count = count_copies(s, pos - ul + 1, ul, 0, pos + 1, false);
    if (bw_count > 1) {
      bw_start = pos - ul - 1;
      bw_len = ul;
      break;
    }
  }
  int total = bw_count;
  if (pos < n - 1) {
    int fw_start = pos + 1, fw_len = 1, fw_count = 0;
    for (int ul = 1; ul <= kMaxRepeatUnit; ++ul) {
      if (pos + ul + 1 > n) break;
      fw_count = count_copies(s, pos + 1, ul, pos + 1, n - pos - 1, true);
      if (fw_count > 1) {
        fw_start = pos + 1;
        fw_len = ul;
        break;
      }
    }
    if (fw_start == bw_start && fw_len == bw_len) {
      total = bw_count + fw_count;
    } else {
      total = count_copies(s, fw_start, fw_len, 0, pos + 1, false) + fw_count;
    }
  }
  return total < kMaxRepeatLen ? total : kMaxRepeatLen;
}

}  // namespace rvc
