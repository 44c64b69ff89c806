// Region-level adapter: what Rovaca's call_pairhmm does around its kernels,
// done on the host around the device engine.
//   read preprocessing        ph/roDV_schedule.cpp:494-649
//   default GOP / GCP         hc/genotype/read_record.cpp:184-210 ('-' = 45, '+'-33 = 10)
//   PCR indel model           ph/pairhmm_internal.cpp:260-314, :391-429; ph/adapter.h:13-15
//   normalize                 ph/pairhmm_internal.cpp:248-258 (ph/adapter.h:23)
//   poorly modelled reads     ph/pairhmm_internal.cpp:22-44, :489-497 (ph/adapter.h:20-22)
//   transpose + erase order   ph/pairhmm_internal.cpp:499-508, :576-623
#include "phmm_adapter.h"
#include "phmm_prep.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <utility>

namespace rvc {
namespace {

// GOP cap per tandem-repeat length 0..20, per model (ph/adapter.h:13-15).
constexpr uint8_t kGopCap[3][21] = {
    {40, 40, 39, 38, 37, 36, 34, 32, 28, 23, 17, 10, 10, 10, 10, 10, 10, 10, 10, 10, 10},   // HOSTILE
    {40, 40, 40, 39, 39, 39, 38, 38, 37, 37, 36, 35, 34, 33, 32, 30, 28, 26, 23, 20, 17},   // AGGRESSIVE
    {40, 40, 40, 40, 39, 39, 39, 39, 39, 38, 38, 38, 37, 37, 37, 36, 36, 35, 34, 33, 33}};  // CONSERVATIVE
}  // namespace

const uint8_t* pcr_gop_caps() { return &kGopCap[0][0]; }

void preprocess_read(const uint8_t* bases, const uint8_t* qual, int len, int mq, int min_q, int pcr, uint8_t* q,
                     uint8_t* gop, uint8_t* gcp) {
  for (int x = 0; x < len; ++x) {
    const uint8_t capped = std::min<uint8_t>(qual[x], (uint8_t)mq);
    q[x] = capped < min_q ? (uint8_t)6 : capped;  // MIN_QUALITY (ph/adapter.h:5)
    gop[x] = (uint8_t)'-';
    gcp[x] = (uint8_t)('+' - 33);
  }
  if (pcr >= RVC_PCR_HOSTILE && pcr <= RVC_PCR_CONSERVATIVE) {
    const uint8_t* cap = kGopCap[pcr - 1];
    // the last base keeps its default
    for (int x = 0; x + 1 < len; ++x) gop[x] = std::min<uint8_t>(gop[x], cap[repeat_length(bases, len, x)]);
  }
}

int postprocess(uint32_t n_reads, uint32_t n_haps, double* read_major, const int32_t* read_len, double* hap_major,
                uint32_t* kept) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (uint32_t r = 0; r < n_reads; ++r) {
    double* row = read_major + (size_t)r * n_haps;
    // the best value as std::max_element finds it: a NaN in front wins
    double best = row[0];
    for (uint32_t h = 1; h < n_haps; ++h)
      if (best < row[h]) best = row[h];
    const double floor_v = best + (-4.5);
    for (uint32_t h = 0; h < n_haps; ++h)
      if (row[h] < floor_v) row[h] = floor_v;
  }
  uint32_t nk = 0;
  for (size_t x = 0; x < (size_t)n_haps * n_reads; ++x) hap_major[x] = nan;
  for (uint32_t r = 0; r < n_reads; ++r) {
    const double* row = read_major + (size_t)r * n_haps;
    double best = -std::numeric_limits<double>::infinity();
    for (uint32_t h = 0; h < n_haps; ++h)
      if (!std::isnan(row[h])) best = std::max(best, row[h]);
    const double min_ll = std::min(2.0, std::ceil(read_len[r] * 0.02)) * (-4.0);
    if (best < min_ll) continue;  // poorly modelled: dropped
    for (uint32_t h = 0; h < n_haps; ++h) hap_major[(size_t)h * n_reads + nk] = row[h];
    kept[nk++] = r;
  }
  return (int)nk;
}

}  // namespace rvc
