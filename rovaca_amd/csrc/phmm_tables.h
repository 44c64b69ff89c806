// Phred -> probability tables of the PairHMM model, built on the host.
#pragma once

#include <cstdint>

namespace rvc {

// Number of match-to-match entries reachable with 7-bit qualities: the
// triangular index max*(max+1)/2 + min for 0 <= min <= max <= 127.
constexpr int kMMEntries = 128 * 129 / 2;

struct HostTables {
  float ph2pr_f[128];
  float mm_f[kMMEntries];
  double ph2pr_d[128];
  double mm_d[kMMEntries];
  uint8_t gop_cap[64];  // PCR indel model: 3 models x 21 repeat lengths
};

// Fills t the way Rovaca's Context<float> / Context<double> do
// (ph/roDV_common/context.h:44-60 jacobian and match-to-match tables in the
// working precision, :98 and :138 ph2pr, :141 / :100 initial constants).
void build_tables(HostTables* t);

}  // namespace rvc
