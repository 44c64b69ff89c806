#pragma once

#include <cstdint>

#include "../../include/rovaca_phmm.h"

namespace rvc {

// Base quality capped by mapping quality and floored to 6 below min_q; gap
// open 45 lowered by the PCR indel model (same array for insertions and
// deletions); gap continuation 10.
void preprocess_read(const uint8_t* bases, const uint8_t* qual, int len, int mq, int min_q, int pcr, uint8_t* q,
                     uint8_t* gop, uint8_t* gcp);

// The PCR indel model's GOP cap per tandem-repeat length 0..20, per model
// (HOSTILE, AGGRESSIVE, CONSERVATIVE): 3 x 21 bytes.
const uint8_t* pcr_gop_caps();

// Normalize each read's row, drop poorly modelled reads and transpose into
// hap-major [n_haps][n_reads] (trailing columns NaN). Returns the number of
// kept reads, listed in kept[].
int postprocess(uint32_t n_reads, uint32_t n_haps, double* read_major, const int32_t* read_len, double* hap_major,
                uint32_t* kept);

}  // namespace rvc
