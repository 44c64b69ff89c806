// Test-only bridge onto Rovaca's own intra-pair PairHMM kernel, compiled from
// the reference sources where they lie (see oracle/Makefile). It exports one C
// function so tests can compare the oracle restatement with the reference's
// actual AVX2/AVX-512 output. Never shipped, never called by the product.
//   ph/roDV_intra/pairhmm_api.cpp:14-35  computeLikelihoodsAVX512 / AVX2
//   ph/roDV_common/common.h:27-62        TestCase, ConvertChar, init_native
#include <cstdint>

#include "roDV_common/common.h"
#include "roDV_intra/pairhmm_api.h"

extern "C" double ref_intra_log10(const uint8_t* hap, int H, const uint8_t* rd, const uint8_t* q, const uint8_t* i,
                                  const uint8_t* d, const uint8_t* c, int R, int use_double) {
  static bool ready = false;
  if (!ready) {
    pairhmm::common::ConvertChar::init();
    ready = true;
  }
  pairhmm::common::init_native();  // FTZ on the calling thread, as the reference does
  pairhmm::common::TestCase tc;
  tc.rslen = (uint32_t)R;
  tc.haplen = (uint32_t)H;
  tc.hap = hap;
  tc.rs = rd;
  tc.q = q;
  tc.i = i;
  tc.d = d;
  tc.c = c;
#if defined(__AVX512F__)
  return pairhmm::intra::computeLikelihoodsAVX512(tc, use_double != 0);
#else
  return pairhmm::intra::computeLikelihoodsAVX2(tc, use_double != 0);
#endif
}

extern "C" int ref_intra_isa(void) {
#if defined(__AVX512F__)
  return 512;
#else
  return 256;
#endif
}
