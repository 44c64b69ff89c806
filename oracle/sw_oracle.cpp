// Smith-Waterman oracle (test infrastructure only): a scalar restatement of
// Rovaca's read-to-haplotype aligner, the ro_sw library behind
// sw_avx_init(0) / avx_function:
//   scores          common/base/lib/sw_avx/sw_avx.c:14-15 (genotype 10/-15/-30/-5,
//                   assembly 200/-150/-260/-11)
//   core choice     common/base/lib/ro_sw/ro_sw.c:26-40 (16-bit scores unless the
//                   read is shorter than 16 or min length x match > INT16_MAX;
//                   the AVX2 build switches below 8 bases, ro_sw.c:10-24, which
//                   gives the same results: under 16 bases no H comes near the
//                   16-bit floor of -8192)
//   recurrence      common/base/lib/ro_sw/sw_core_template.h:113-330
//   best end cell   sw_core_template.h:332-348 (last row, then last column,
//                   one anti-diagonal at a time)
//   traceback/CIGAR common/base/lib/ro_sw/sw_cigar.c:20-171
// Rows i = 1..len1 are haplotype bases (seq1), columns j = 1..len2 read bases
// (seq2); row 0 and column 0 hold H = 0, so the alignment may start anywhere
// on the haplotype and the read's unaligned ends become soft clips.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace {
constexpr int kMaxSeqLen = 2048;  // sw_avx.c:18
// backtrack codes (smithwaterman_common.h:40-49)
constexpr int kMatch = 0, kIns = 1, kDel = 2, kInsExt = 4, kDelExt = 8, kSoftClip = 9;
// BAM CIGAR operations (htslib sam.h)
constexpr uint32_t kOpM = 0, kOpI = 1, kOpD = 2, kOpS = 4;
}  // namespace

extern "C" int orc_sw_align(const uint8_t* s1, int n1, const uint8_t* s2, int n2, int match, int mismatch, int open,
                            int extend, uint32_t* cigar, int max_cigar, int* offset, int* score) {
  if (std::max(n1, n2) >= kMaxSeqLen) {
    *offset = 1;
    return -1;
  }
  // 16-bit core: H floored at INT16_MIN/4, boundary gaps at INT16_MIN/2;
  // 32-bit core: -100000000 and INT32_MIN/2 (sw_internal.h:19-22). The
  // arithmetic is exact in both (no 16-bit sum can leave the range: every gap
  // score is at least the floor plus one open).
  const bool use32 = n2 < 16 || match <= 0 || (long long)std::min(n1, n2) * match > INT16_MAX;
  const int low = use32 ? INT32_MIN / 2 : INT16_MIN / 2;
  const int cut = use32 ? -100000000 : INT16_MIN / 4;
  const int W = n2 + 1;
  std::vector<int> H((size_t)(n1 + 1) * W, 0), E((size_t)(n1 + 1) * W, low), F((size_t)(n1 + 1) * W, low);
  std::vector<uint8_t> bt((size_t)(n1 + 1) * W, 0);
  for (int i = 1; i <= n1; ++i)
    for (int j = 1; j <= n2; ++j) {
      const size_t c = (size_t)i * W + j;
      const int oh = H[c - 1] + open, eh = E[c - 1] + extend;
      const int e = std::max(oh, eh);
      const int ins_ext = (oh > eh) ? 0 : kInsExt;
      const int ov = H[c - W] + open, ev = F[c - W] + extend;
      const int f = std::max(ev, ov);
      const int del_ext = (ov > ev) ? 0 : kDelExt;
      const int m = H[c - W - 1] + (s1[i - 1] == s2[j - 1] ? match : mismatch);
      int h = std::max(cut, m);
      int b = (e > h) ? kIns : kMatch;
      h = std::max(h, e);
      if (f > h) b = kDel;
      h = std::max(h, f);
      E[c] = e;
      F[c] = f;
      H[c] = h;
      bt[c] = (uint8_t)(b | ins_ext | del_ext);
    }
  // best end cell: anti-diagonals d = 1 .. n1 + n2; on each, the last-row
  // cell (n1, d - n1) and then the last-column cell (d - n2, n2)
  int best = use32 ? INT32_MIN : INT16_MIN, bi = 0, bj = 0;
  for (int d = 1; d <= n1 + n2; ++d) {
    if (d >= n1 + 1) {
      const int i = n1, j = d - n1, v = H[(size_t)i * W + j];
      if (best < v || (best == v && std::abs(i - j) < std::abs(bi - bj))) { best = v; bi = i; bj = j; }
    }
    if (d >= n2 + 1) {
      const int i = d - n2, j = n2, v = H[(size_t)i * W + j];
      if (best < v || (best == v && (bj == n2 || std::abs(i - j) <= std::abs(bi - bj)))) { best = v; bi = i; bj = j; }
    }
  }
  *score = best;
  // traceback from the best cell: runs collected end to start
  std::vector<std::pair<int, int>> runs;
  int i = bi, j = bj;
  if (j < n2) runs.push_back({kSoftClip, n2 - j});
  int state = 0;
  while (i > 0 && j > 0) {
    const int b = bt[(size_t)i * W + j];
    if (state == kInsExt) {
      --j;
      runs.back().second++;
      state = b & kInsExt;
    } else if (state == kDelExt) {
      --i;
      runs.back().second++;
      state = b & kDelExt;
    } else {
      switch (b & 3) {
        case kMatch: --i; --j; runs.push_back({kMatch, 1}); state = 0; break;
        case kIns: --j; runs.push_back({kIns, 1}); state = b & kInsExt; break;
        case kDel: --i; runs.push_back({kDel, 1}); state = b & kDelExt; break;
      }
    }
  }
  if (j > 0) runs.push_back({kSoftClip, j});
  *offset = i;
  // merge equal neighbours, then emit start to end
  std::vector<std::pair<int, int>> merged;
  for (const auto& r : runs) {
    if (!merged.empty() && merged.back().first == r.first) merged.back().second += r.second;
    else merged.push_back(r);
  }
  int n = 0;
  for (int x = (int)merged.size() - 1; x >= 0 && n < max_cigar; --x) {
    if (merged[x].second <= 0) continue;
    const int st = merged[x].first;
    const uint32_t op = st == kMatch ? kOpM : st == kIns ? kOpI : st == kDel ? kOpD : kOpS;
    cigar[n++] = ((uint32_t)merged[x].second << 4) | op;
  }
  return n;
}
