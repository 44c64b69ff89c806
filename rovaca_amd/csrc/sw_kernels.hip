// Smith-Waterman of reads against haplotypes on gfx950: Rovaca's
// read-to-best-haplotype aligner (the ro_sw library behind sw_avx_init /
// avx_function, common/base/lib/ro_sw/sw_core_template.h:110-361 and
// sw_cigar.c:20-171), batched.
//
// Matrix: rows j = 1..n2 are read bases, columns i = 1..n1 haplotype bases
// (the reference's i and j, transposed for the sweep). Per cell
//   E = max(H[i][j-1] + open, E[i][j-1] + extend)   (insertion, down a column)
//   F = max(F[i-1][j] + extend, H[i-1][j] + open)   (deletion, along a row)
//   H = max(max(cut, H[i-1][j-1] + s(hap[i], read[j])), E, F)
// with the reference's tie order (E replaces M only when greater, F replaces
// both only when greater) and a backtrack nibble: bit 0 E > M (insertion),
// bit 1 F > max(M, E) (deletion, wins over bit 0), bit 2 E was opened here
// (not extended), bit 3 F was opened here. Row 0 and column 0 hold H = 0 and
// no open gap.
//
// The sweep is the PairHMM one: a segment of W lanes takes one pair, lane k
// holds read rows k*RPL .. k*RPL+RPL-1 and at step s computes haplotype
// column s-k+1 of all of them; H and E of a lane's last row move one lane down
// with DPP, F stays in the lane. Rows past the read carry H of the read's last
// row down unchanged (E open 0, no match, no deletion), so the segment's last
// row holds H[i][n2] for the best-end search. Backtrack nibbles are stored per
// step in lane order (coalesced, 16 bits per 4 rows); the traceback kernel
// maps (i, j) back. The packed int16 sweep (k_sw_fill16, below) runs the same
// recurrence two pairs per lane.
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "sw_internal.h"

namespace rvc {
namespace {

constexpr int kHapRow = 2176;  // LDS bytes per segment: 64 zeros, <= 2047 bases, drain columns

template <int W>
__device__ __forceinline__ int down0(int v, int k) {
  if constexpr (W == 16) {
    return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
  } else if constexpr (W == 64) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
  } else {
    const int t = __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
    return k == 0 ? 0 : t;
  }
}

template <int W>
__device__ __forceinline__ int down_or(int v, int old, int k) {
  if constexpr (W == 16) {
    return __builtin_amdgcn_update_dpp(old, v, 0x111, 0xf, 0xf, false);
  } else {
    const int t = __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false);
    return k == 0 ? old : t;
  }
}

// One step of a lane over its RPL rows at haplotype base hc.
// uH, uE: H and E of the row above in this column; dH: H of the row above in
// the previous column.
template <int RPL>
__device__ __forceinline__ void sw_step(int (&Hp)[RPL], int (&F)[RPL], int (&Hn)[RPL], int& lastE,
                                        uint32_t (&bt)[sw_groups(RPL)],
                                        int uH, int uE, int dH, int hc, const int (&rc)[RPL], const int (&oE)[RPL],
                                        const int (&oF)[RPL], const int (&mis)[RPL], int match, int ext, int cut) {
  int aH = uH, aE = uE, dg = dH;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int oh = aH + oE[r], eh = aE + ext;
    const int e = max(oh, eh);
    const int ov = Hp[r] + oF[r], ev = F[r] + ext;
    const int f = max(ev, ov);
    const int m = dg + (hc == rc[r] ? match : mis[r]);
    const int h0 = max(cut, m);
    // H = max(M, E, F); E comes from the row above in this column, so only
    // the last max waits for it
    const int h = max(max(h0, f), e);
    const int h1 = max(h0, e);
    // the four decisions as sign bits of differences (all well inside int32),
    // shifted into the group's 16 bits: row q of a 4-row group ends at bits
    // 4 * (3 - q) .. 4 * (3 - q) + 3
    uint32_t acc = (r & 3) == 0 ? 0u : bt[r >> 2];
    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(ev - ov), 31);  // bit 3: F opened (ov > ev)
    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(eh - oh), 31);  // bit 2: E opened (oh > eh)
    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(h1 - f), 31);   // bit 1: f > max(M, E)
    acc = __builtin_amdgcn_alignbit(acc, (uint32_t)(h0 - e), 31);   // bit 0: e > M
    bt[r >> 2] = acc;
    dg = Hp[r];
    Hn[r] = h;
    F[r] = f;
    aH = h;
    aE = e;
  }
  // a short last group: its nibbles to the places of a full group's
  if constexpr (RPL % 4 != 0) bt[RPL / 4] <<= 4 * (4 - RPL % 4);
  lastE = aE;
}

template <int W, int RPL>
__global__ __launch_bounds__(64) void k_sw_fill(SwArgs a) {
  constexpr int NSEG = 64 / W;
  __shared__ uint8_t hb[NSEG][kHapRow];
  const int lane = threadIdx.x;
  const int seg = lane / W, k = lane % W;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.head, (uint32_t)NSEG);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.n_list) break;
    const bool active = t + (uint32_t)seg < a.n_list;
    const uint32_t p = a.list[active ? t + seg : t];
    const SwPair pr = a.pairs[p];
    const int n1 = (int)pr.hap_len, n2 = (int)pr.read_len;
    // the wave sweeps to its longest haplotype
    int nmax = n1;
#pragma unroll
    for (int g = 0; g < NSEG; ++g) {
      const int o = __shfl(n1, g * W);
      nmax = o > nmax ? o : nmax;
    }
    __syncthreads();
    for (int x = k; x < kHapRow; x += W) {
      const int i = x - 64;
      hb[seg][x] = (i >= 0 && i < n1) ? a.haps[pr.hap_off + i] : 0;
    }
    __syncthreads();
    int rc[RPL], oE[RPL], oF[RPL], mis[RPL], Hp[RPL], F[RPL], Hn[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int j = k * RPL + r;
      const bool real = j < n2;
      rc[r] = real ? (int)a.reads[pr.read_off + j] : 0x1ff;
      oE[r] = real ? a.open : 0;
      oF[r] = real ? a.open : kSwBig;
      mis[r] = real ? a.mismatch : kSwBig;
      Hp[r] = 0;
      F[r] = kSwLow;
    }
    // the reference's 16-bit core floors H at INT16_MIN / 4 (ro_sw.c:26-40,
    // sw_internal.h:20-23); its 32-bit core at -100000000
    const bool use32 = n2 < 16 || a.match <= 0 || (long long)(n1 < n2 ? n1 : n2) * a.match > 32767;
    const int cut = use32 ? -100000000 : -8192;
    constexpr int G = sw_groups(RPL);
    uint16_t* btp = reinterpret_cast<uint16_t*>(a.bt + pr.bt_off) + k * G;
    int32_t* hrow = a.hrow + pr.hrow_off;
    int32_t* hcol = a.hcol + pr.hcol_off;
    int uH = 0, uE = kSwLow, dH = 0;
    const int own_steps = n1 + W - 1;
    const int nsteps = nmax + W - 1;
    int hc = hb[seg][64 - k];  // this step's haplotype base, read a step ahead
    auto step = [&](int s, auto fill) {
      const int col = s - k + 1;  // haplotype column of this lane
      const int hc_next = hb[seg][64 + s + 1 - k];
      uint32_t bt[G];
      int lastE;
      sw_step<RPL>(Hp, F, Hn, lastE, bt, uH, uE, dH, hc, rc, oE, oF, mis, a.match, a.extend, cut);
      if constexpr (decltype(fill)::value) {
        if (col < 1) {
          // before the first column: keep column 0 (H = 0, no deletion)
#pragma unroll
          for (int r = 0; r < RPL; ++r) { Hn[r] = 0; F[r] = kSwLow; }
        }
      }
      if (active && s < own_steps) {
        uint16_t* dst = btp + (size_t)s * (W * G);
#pragma unroll
        for (int q = 0; q < G; ++q) dst[q] = (uint16_t)bt[q];
        if (col == n1) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) hcol[k * RPL + r] = Hn[r];
        }
        if (k == W - 1 && col >= 1 && col <= n1) hrow[col - 1] = Hn[RPL - 1];
      }
      dH = uH;
      uH = down0<W>(Hn[RPL - 1], k);
      uE = down_or<W>(lastE, kSwLow, k);
#pragma unroll
      for (int r = 0; r < RPL; ++r) Hp[r] = Hn[r];
      hc = hc_next;
    };
    // the first W - 1 steps start lanes one by one; after them every lane
    // is past column 0
    int s = 0;
    for (; s < W - 1 && s < nsteps; ++s) step(s, std::true_type{});
    for (; s < nsteps; ++s) step(s, std::false_type{});
  }
}

// ---- the packed int16 sweep: two pairs per lane ---------------------------
// The same recurrence, tie order and backtrack nibbles on pairs that
// sw_narrow() admits, with pair A in the low and pair B in the high 16 bits
// of every register (v_pk_* operations). The match score is
// (dg + match) + min(hc ^ rc, 1) * (mismatch - match), byte equality as in
// the int32 sweep; rows past a read hold rc = 0x1ff (never equal), mismatch
// and deletion open kSwBig16, insertion open 0. Each 4-row group leaves one
// 32-bit word per step, A's 16 bits low and B's high, with its own bit order
// (see sw_step16; k_sw_trace reads both orders).
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 pmax(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ uint32_t bits(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s16x2 halves(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ s16x2 splat(int v) { return s16x2{(short)v, (short)v}; }

template <int RPL>
__device__ __forceinline__ void sw_step16(s16x2 (&Hp)[RPL], s16x2 (&F)[RPL], s16x2 (&Hn)[RPL], s16x2& lastE,
                                          uint32_t (&bt)[sw_groups(RPL)], s16x2 uH, s16x2 uE, s16x2 dH, u16x2 hc,
                                          const u16x2 (&rc)[RPL], const s16x2 (&oE)[RPL], const s16x2 (&oF)[RPL],
                                          const s16x2 (&dl)[RPL], s16x2 match, s16x2 ext, s16x2 cut, u16x2 one) {
  s16x2 aH = uH, aE = uE, dg = dH;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const s16x2 oh = aH + oE[r], eh = aE + ext;
    const s16x2 e = pmax(oh, eh);
    const s16x2 ov = Hp[r] + oF[r], ev = F[r] + ext;
    const s16x2 f = pmax(ev, ov);
    const s16x2 t = __builtin_bit_cast(s16x2, __builtin_elementwise_min(hc ^ rc[r], one));
    const s16x2 m = t * dl[r] + (dg + match);
    const s16x2 h0 = pmax(cut, m);
    const s16x2 h1 = pmax(h0, e);
    const s16x2 h = pmax(h1, f);
    // the high bytes of the four differences (signs at bit 7 of each byte)
    // gathered two at a time, then shifted into the group's word: in each
    // half, row q of the group ends with F opened at bit 2q, deletion at
    // 2q + 1, E opened at 8 + 2q, insertion at 9 + 2q
    const uint32_t p01 = __builtin_amdgcn_perm(bits(eh - oh), bits(ev - ov), 0x07030501u);
    const uint32_t p23 = __builtin_amdgcn_perm(bits(h0 - e), bits(h1 - f), 0x07030501u);
    uint32_t acc = (r & 3) == 0 ? 0u : bt[r >> 2];
    acc = (p01 & 0x80808080u) | ((acc >> 1) & 0x7f7f7f7fu);
    acc = (p23 & 0x80808080u) | ((acc >> 1) & 0x7f7f7f7fu);
    bt[r >> 2] = acc;
    dg = Hp[r];
    Hn[r] = h;
    F[r] = f;
    aH = h;
    aE = e;
  }
  // a short last group: its bits down to the places of a full group's
  if constexpr (RPL % 4 != 0) {
    constexpr int sh = 2 * (4 - RPL % 4);
    constexpr uint32_t keep = (0xffu >> sh) * 0x01010101u;
    bt[RPL / 4] = (bt[RPL / 4] >> sh) & keep;
  }
  lastE = aE;
}

template <int W, int RPL>
__global__ __launch_bounds__(64) void k_sw_fill16(SwArgs a) {
  constexpr int NSEG = 64 / W;
  // haplotype bases of the duo, A in the low and B in the high byte
  __shared__ uint16_t hb[NSEG][kHapRow];
  const int lane = threadIdx.x;
  const int seg = lane / W, k = lane % W;
  const s16x2 low = splat(kSwLow16);
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.head, (uint32_t)NSEG);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.n_list) break;
    const bool active = t + (uint32_t)seg < a.n_list;
    const uint32_t d = active ? t + seg : t;
    const uint32_t pa = a.list[2 * d], pb0 = a.list[2 * d + 1];
    const bool has_b = active && pb0 != kSwNoPair;
    const uint32_t pb = pb0 != kSwNoPair ? pb0 : pa;
    const SwPair A = a.pairs[pa], B = a.pairs[pb];
    const int n1a = (int)A.hap_len, n1b = (int)B.hap_len;
    const int n2a = (int)A.read_len, n2b = (int)B.read_len;
    const int n1 = n1a > n1b ? n1a : n1b;
    int nmax = n1;
#pragma unroll
    for (int g = 0; g < NSEG; ++g) {
      const int o = __shfl(n1, g * W);
      nmax = o > nmax ? o : nmax;
    }
    __syncthreads();
    for (int x = k; x < kHapRow; x += W) {
      const int i = x - 64;
      const uint32_t ba = (i >= 0 && i < n1a) ? a.haps[A.hap_off + i] : 0u;
      const uint32_t bb = (i >= 0 && i < n1b) ? a.haps[B.hap_off + i] : 0u;
      hb[seg][x] = (uint16_t)(ba | bb << 8);
    }
    __syncthreads();
    u16x2 rc[RPL];
    s16x2 oE[RPL], oF[RPL], dl[RPL], Hp[RPL], F[RPL], Hn[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int j = k * RPL + r;
      const bool ra = j < n2a, rb = j < n2b;
      rc[r] = u16x2{(unsigned short)(ra ? a.reads[A.read_off + j] : 0x1ff),
                    (unsigned short)(rb ? a.reads[B.read_off + j] : 0x1ff)};
      oE[r] = s16x2{(short)(ra ? a.open : 0), (short)(rb ? a.open : 0)};
      oF[r] = s16x2{(short)(ra ? a.open : kSwBig16), (short)(rb ? a.open : kSwBig16)};
      dl[r] = s16x2{(short)((ra ? a.mismatch : kSwBig16) - a.match), (short)((rb ? a.mismatch : kSwBig16) - a.match)};
      Hp[r] = splat(0);
      F[r] = low;
    }
    const s16x2 match = splat(a.match), ext = splat(a.extend), cut = splat(-8192);
    // 1 in both halves, kept opaque: min(x, 1) stays one v_pk_min_u16 rather
    // than a compare and select per half
    uint32_t one_bits = 0x00010001u;
    asm volatile("" : "+v"(one_bits));
    const u16x2 one = __builtin_bit_cast(u16x2, one_bits);
    constexpr int G = sw_groups(RPL);
    uint32_t* btp = reinterpret_cast<uint32_t*>(a.bt + A.bt_off) + k * G;
    int32_t* hrow_a = a.hrow + A.hrow_off;
    int32_t* hcol_a = a.hcol + A.hcol_off;
    int32_t* hrow_b = a.hrow + B.hrow_off;
    int32_t* hcol_b = a.hcol + B.hcol_off;
    s16x2 uH = splat(0), uE = low, dH = splat(0);
    const int own_steps = n1 + W - 1;
    const int nsteps = nmax + W - 1;
    // this step's haplotype bases (read a step ahead), one per half
    uint32_t hc = __builtin_amdgcn_perm(0u, (uint32_t)hb[seg][64 - k], 0x0c010c00u);
    auto step = [&](int s, auto fill) {
      const int col = s - k + 1;
      const uint32_t hc_next = __builtin_amdgcn_perm(0u, (uint32_t)hb[seg][64 + s + 1 - k], 0x0c010c00u);
      uint32_t bt[G];
      s16x2 lastE;
      sw_step16<RPL>(Hp, F, Hn, lastE, bt, uH, uE, dH, __builtin_bit_cast(u16x2, hc), rc, oE, oF, dl, match, ext, cut, one);
      if constexpr (decltype(fill)::value) {
        if (col < 1) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) { Hn[r] = splat(0); F[r] = low; }
        }
      }
      if (active && s < own_steps) {
        uint32_t* dst = btp + (size_t)s * (W * G);
#pragma unroll
        for (int q = 0; q < G; ++q) dst[q] = bt[q];
        if (col == n1a) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) hcol_a[k * RPL + r] = Hn[r].x;
        }
        if (has_b && col == n1b) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) hcol_b[k * RPL + r] = Hn[r].y;
        }
        if (k == W - 1 && col >= 1) {
          if (col <= n1a) hrow_a[col - 1] = Hn[RPL - 1].x;
          if (has_b && col <= n1b) hrow_b[col - 1] = Hn[RPL - 1].y;
        }
      }
      dH = uH;
      uH = halves((uint32_t)down0<W>((int)bits(Hn[RPL - 1]), k));
      uE = halves((uint32_t)down_or<W>((int)bits(lastE), (int)bits(low), k));
#pragma unroll
      for (int r = 0; r < RPL; ++r) Hp[r] = Hn[r];
      hc = hc_next;
    };
    int s = 0;
    for (; s < W - 1 && s < nsteps; ++s) step(s, std::true_type{});
    for (; s < nsteps; ++s) step(s, std::false_type{});
  }
}

// One wave per pair: the best end cell over the last haplotype column
// (H[n1][j]) and the last read row (H[i][n2]), scanned one anti-diagonal
// d = i + j at a time, column candidate before row candidate, with the
// reference's replacement rules (sw_core_template.h:332-348): a candidate
// replaces the best if greater, or if equal and closer to the diagonal
// (a row candidate also when equally close, or when the best is on the last
// row). Only candidates equal to the maximum can be the result, and the first
// of them in scan order replaces whatever came before, so the wave takes the
// maximum with coalesced loads, then applies the rules to the equal ones in
// scan order.
__global__ __launch_bounds__(256) void k_sw_best(SwArgs a) {
  const uint32_t x = blockIdx.x * 4 + threadIdx.x / 64;
  const int l = threadIdx.x % 64;
  if (x >= a.n_list) return;  // whole waves
  const uint32_t p = a.list[x];
  if (p == kSwNoPair) return;
  const SwPair pr = a.pairs[p];
  const int n1 = (int)pr.hap_len, n2 = (int)pr.read_len;
  if (n1 >= kSwMaxLen || n2 >= kSwMaxLen || n1 == 0 || n2 == 0) return;
  const int32_t* hrow = a.hrow + pr.hrow_off;  // H[i][n2], i = 1..n1
  const int32_t* hcol = a.hcol + pr.hcol_off;  // H[n1][j], j = 1..n2
  const int d_lo = (n1 < n2 ? n1 : n2) + 1, d_hi = n1 + n2;
  int vmax = INT32_MIN;
  for (int d0 = d_lo; d0 <= d_hi; d0 += 64) {
    const int d = d0 + l;
    if (d <= d_hi) {
      if (d >= n1 + 1) vmax = max(vmax, hcol[d - n1 - 1]);
      if (d >= n2 + 1) vmax = max(vmax, hrow[d - n2 - 1]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) vmax = max(vmax, __shfl_xor(vmax, o));
  int bi = 0, bj = 0;
  bool have = false;
  for (int d0 = d_lo; d0 <= d_hi; d0 += 64) {
    const int d = d0 + l;
    const bool c = d <= d_hi && d >= n1 + 1 && hcol[d - n1 - 1] == vmax;
    const bool r = d <= d_hi && d >= n2 + 1 && hrow[d - n2 - 1] == vmax;
    uint64_t mask = __ballot(c || r);
    while (mask) {
      const int src = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int dd = d0 + src;
      if (__shfl(c ? 1 : 0, src)) {
        const int i = n1, j = dd - n1;
        if (!have || abs(i - j) < abs(bi - bj)) { bi = i; bj = j; have = true; }
      }
      if (__shfl(r ? 1 : 0, src)) {
        const int i = dd - n2, j = n2;
        if (!have || bj == n2 || abs(i - j) <= abs(bi - bj)) { bi = i; bj = j; have = true; }
      }
    }
  }
  if (l == 0) {
    a.score[p] = vmax;
    a.best_ij[2 * p] = bi;
    a.best_ij[2 * p + 1] = bj;
  }
}

// One thread per pair: the traceback from the best end cell into CIGAR runs,
// written start to end.
__global__ __launch_bounds__(64) void k_sw_trace(SwArgs a) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n_list) return;
  const uint32_t p = a.list[x];
  if (p == kSwNoPair) return;
  const SwPair pr = a.pairs[p];
  const int n2 = (int)pr.read_len;
  uint32_t* out = a.cigar + (size_t)p * a.max_cigar;
  const int bi = a.best_ij[2 * p], bj = a.best_ij[2 * p + 1];
  const int rpl = (int)pr.rpl, gpl = sw_groups(rpl), wg = (int)pr.width * gpl;
  // 16-bit words: the pair's own, or every other one of a duo's 32-bit words
  const uint16_t* bt = reinterpret_cast<const uint16_t*>(a.bt + pr.bt_off) + pr.bt_half;
  const int bt_lanes = pr.bt_lanes;
  // runs end to start, merged as they come
  int n = 0, cur_op = -1;
  uint32_t cur_len = 0;
  bool overflow = false;
  auto push = [&](int op, uint32_t len) {
    if (op == cur_op) { cur_len += len; return; }
    if (cur_op >= 0) {
      if (n < (int)a.max_cigar) out[n] = (cur_len << 4) | (uint32_t)cur_op;
      else overflow = true;
      ++n;
    }
    cur_op = op;
    cur_len = len;
  };
  int i = bi, j = bj;
  if (j < n2) push(kSwOpS, (uint32_t)(n2 - j));
  int state = 0;
  // one move of the walk from the current cell, given its backtrack word
  auto move = [&](uint32_t w16, int q) {
    // bit 0 insertion, 1 deletion, 2 E opened, 3 F opened
    const int v = bt_lanes == 1 ? (int)(w16 >> (4 * (3 - q))) & 15
                                : (int)(((w16 >> (9 + 2 * q)) & 1) | ((w16 >> (2 * q)) & 2) |
                                        ((w16 >> (6 + 2 * q)) & 4) | ((w16 << 3 >> (2 * q)) & 8));
    // state 4: inside an insertion, 8: inside a deletion (the gap goes on
    // where this cell's E / F was extended, not opened)
    const int ins_ext = (v & 4) ? 0 : 4, del_ext = (v & 8) ? 0 : 8;
    if (state == 4) {
      --j;
      push(kSwOpI, 1);
      state = ins_ext;
    } else if (state == 8) {
      --i;
      push(kSwOpD, 1);
      state = del_ext;
    } else if (v & 2) {
      --i;
      push(kSwOpD, 1);
      state = del_ext;
    } else if (v & 1) {
      --j;
      push(kSwOpI, 1);
      state = ins_ext;
    } else {
      --i;
      --j;
      push(kSwOpM, 1);
      state = 0;
    }
  };
  // The walk is a chain of dependent loads, one backtrack word per cell. With
  // a.trace_run > 1 the words of the next a.trace_run cells along the
  // diagonal (a run of matches, the common move) are loaded together, so such
  // a run waits for one load per a.trace_run cells; the walk takes them while
  // its moves stay on that diagonal and loads again from where it leaves it.
  // Decisions are the same, cell by cell. That pays where few walks run at
  // once (latency-bound); with many, the loads the walk leaves unused cost
  // more than the waits they save, and the host passes 1 (a.trace_run is the
  // same for the whole launch).
  constexpr int kRun = 8;
  const int run = min(kRun, max(1, a.trace_run));
  if (run == 1) {
    while (i > 0 && j > 0) {
      const int kk = (j - 1) / rpl;  // the lane of read row j
      const int rr = (j - 1) - kk * rpl;
      const size_t word = (size_t)(i - 1 + kk) * wg + (size_t)(kk * gpl + (rr >> 2));
      move(bt[word * bt_lanes], rr & 3);
    }
  }
  while (i > 0 && j > 0) {
    const int i0 = i, j0 = j;
    const int nl = min(run, min(i, j));  // cells (i0 - t, j0 - t), t < nl, are inside the matrix
    uint32_t wv[kRun];
    int qv[kRun];
#pragma unroll
    for (int t = 0; t < kRun; ++t) {
      wv[t] = 0;
      qv[t] = 0;
      if (t < nl) {
        const int jj = j0 - t;
        const int kk = (jj - 1) / rpl;  // the lane of read row jj
        const int rr = (jj - 1) - kk * rpl;
        const size_t word = (size_t)(i0 - t - 1 + kk) * wg + (size_t)(kk * gpl + (rr >> 2));
        wv[t] = bt[word * bt_lanes];
        qv[t] = rr & 3;
      }
    }
#pragma unroll
    for (int t = 0; t < kRun; ++t) {
      if (t >= nl || i != i0 - t || j != j0 - t) continue;  // off the loaded diagonal
      move(wv[t], qv[t]);
    }
  }
  if (j > 0) push(kSwOpS, (uint32_t)j);
  push(-2, 0);  // flush the last run
  if (overflow) {
    a.count[p] = -2;
    a.offset[p] = i;
    return;
  }
  for (int x = 0, y = n - 1; x < y; ++x, --y) {
    const uint32_t tmp = out[x];
    out[x] = out[y];
    out[y] = tmp;
  }
  a.count[p] = n;
  a.offset[p] = i;
}

using SwFn = void (*)(SwArgs);
template <int... S>
constexpr std::array<SwFn, kSwShapes> fill_table(std::integer_sequence<int, S...>) {
  return {&k_sw_fill<kSwW[S], kSwR[S]>...};
}
template <int... S>
constexpr std::array<SwFn, kSwShapes> fill16_table(std::integer_sequence<int, S...>) {
  // no packed sweep of the 2048-row shape
  return {(S == kSwLongShape ? nullptr : &k_sw_fill16<kSwW[S], S == kSwLongShape ? 4 : kSwR[S]>)...};
}
constexpr std::array<SwFn, kSwShapes> kSwFill = fill_table(std::make_integer_sequence<int, kSwShapes>{});
constexpr std::array<SwFn, kSwShapes> kSwFill16 = fill16_table(std::make_integer_sequence<int, kSwShapes>{});

// Occupancy per shape, cached; asked from several threads at once (each
// slot is atomic, and racing threads store the same value).
int blocks_per_cu(const std::array<SwFn, kSwShapes>& table, std::atomic<int>* cache, int shape) {
  if (shape < 0 || shape >= kSwShapes || !table[shape]) return 1;
  int n = cache[shape].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, table[shape], 64, 0) != hipSuccess || n < 1) n = 1;
    cache[shape].store(n, std::memory_order_relaxed);
  }
  return n;
}
}  // namespace

hipError_t launch_sw_fill(int shape, int grid, hipStream_t stream, const SwArgs& a) {
  if (shape < 0 || shape >= kSwShapes) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kSwFill[shape], dim3(grid), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_sw_trace(hipStream_t stream, const SwArgs& a) {
  if (a.n_list == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sw_best, dim3((a.n_list + 3) / 4), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(k_sw_trace, dim3((a.n_list + 63) / 64), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_sw_fill16(int shape, int grid, hipStream_t stream, const SwArgs& a) {
  if (!sw_packed_shape(shape)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kSwFill16[shape], dim3(grid), dim3(64), 0, stream, a);
  return hipGetLastError();
}

int sw_blocks_per_cu(int shape) {
  static std::atomic<int> cache[kSwShapes] = {};
  return blocks_per_cu(kSwFill, cache, shape);
}

int sw16_blocks_per_cu(int shape) {
  static std::atomic<int> cache[kSwShapes] = {};
  return blocks_per_cu(kSwFill16, cache, shape);
}

}  // namespace rvc
