// PairHMM forward kernels for gfx950 (CDNA4).
//
// Layout of the work inside one wavefront (64 lanes): the wave is cut into
// 64/W segments of W lanes; each segment sweeps one read x haplotype matrix.
// Lane k of a segment owns read rows k*RPL .. k*RPL+RPL-1 (registers), and at
// step s it computes column j = s - k of all its rows, so the segment works
// on a skewed band of W anti-diagonals at once. The three values of a lane's
// last row (M, I, D) move one lane down per step through DPP (row_shr:1 for
// W = 16, wave_shr:1 for W = 32/64), which feeds the first row of the next
// lane one step later; lane 0 of a segment takes the row -1 boundary
// (M = I = 0, D = C/H) from the DPP 'old' operand. The haplotype is staged in
// LDS as one code byte per base (compare kernels) or as 2-bit codes
// (prior-table kernels); lane k reads column (s - k) each step. The
// prior-table kernels (FP32 and FP64) move two values instead: the lane above
// mixes its last row with the receiving lane's row-0 coefficients
// (sweep_step_xy).
//
// Rows past the end of a read are padding rows that carry the sum
// M[R-1][j] + I[R-1][j] down unchanged (prior 0, I <- M + I), so the result
// always comes out of the segment's last lane and last row register (the
// 64-lane FP64 bin also takes reads that fill every row: there the last row's
// M + I is summed directly).
//
// Recurrence and boundaries follow Rovaca's kernels:
//   ph/roDV_inter/pairhmm_inter.cpp:16-102 (per-row parameters), :170-177
//   (M/I/D update), :230-244 (final-row sum), :256-271 (row -1 = C/H);
//   ph/roDV_intra/pairhmm_impl.cpp:13-44 (N matches every base).
// FP32 uses C = 2^120 and hands pairs whose sum is below 1e-28 to the FP64
// kernel (C = 2^1020): ph/roDV_schedule.cpp:35,702-713; ph/roDV_intra/pairhmm_api.cpp:14-23.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "phmm_internal.h"
#include "phmm_prep.h"

namespace rvc {
namespace {

template <typename T>
struct Num;
template <>
struct Num<float> {
  __device__ static float C() { return 0x1p120f; }
};
template <>
struct Num<double> {
  __device__ static double C() { return 0x1p1020; }
};

// Segment and position of a lane. W = 8 interleaves two segments in each
// 16-lane DPP row (even and odd lanes), so that row_shr:2 moves a value one
// position down inside its segment and the row's lanes 0 and 1 are the two
// segments' first positions; the other widths use consecutive lanes.
template <int W>
__device__ __forceinline__ int seg_of(int lane) {
  if constexpr (W == 8) return ((lane >> 4) << 1) | (lane & 1);
  else return lane / W;
}
template <int W>
__device__ __forceinline__ int pos_of(int lane) {
  if constexpr (W == 8) return (lane & 15) >> 1;
  else return lane % W;
}

// DPP control of the one-position move: row_shr:2 (W = 8), row_shr:1 (16),
// wave_shr:1 (32, 64).
template <int W>
constexpr int kDownCtrl = W == 8 ? 0x112 : W == 16 ? 0x111 : 0x138;

// Move v one lane down inside each W-lane segment; lane 0 of a segment
// receives 0 (DPP bound control: no register has to be cleared first).
template <int W>
__device__ __forceinline__ float lane_down0(float v, int k) {
  if constexpr (W == 8) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xf, 0xf, true));
  } else if constexpr (W == 16) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
  } else if constexpr (W == 64) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
  } else {
    const float t = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
    return k == 0 ? 0.f : t;
  }
}

// The same, with lane 0 of a segment receiving 'old'.
template <int W>
__device__ __forceinline__ float lane_down(float v, float old, int k) {
  if constexpr (W == 8) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x112, 0xf, 0xf, false));
  } else if constexpr (W == 16) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x111, 0xf, 0xf, false));
  } else if constexpr (W == 64) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
  } else {
    const float t =
        __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
    return k == 0 ? old : t;
  }
}


template <int W>
__device__ __forceinline__ double lane_down(double v, double old, int k) {
  const long long vb = __double_as_longlong(v), ob = __double_as_longlong(old);
  const int vlo = (int)(vb & 0xffffffffLL), vhi = (int)(vb >> 32);
  const int olo = (int)(ob & 0xffffffffLL), ohi = (int)(ob >> 32);
  int lo, hi;
  if constexpr (W == 16) {
    lo = __builtin_amdgcn_update_dpp(olo, vlo, 0x111, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(ohi, vhi, 0x111, 0xf, 0xf, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(olo, vlo, 0x138, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(ohi, vhi, 0x138, 0xf, 0xf, false);
  }
  double r = __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
  if constexpr (W == 32) r = (k == 0) ? old : r;
  return r;
}

// Double versions: two 32-bit DPP moves per value.
template <int W>
__device__ __forceinline__ double lane_down0(double v, int k) {
  const long long vb = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(vb & 0xffffffffLL), kDownCtrl<W>, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(vb >> 32), kDownCtrl<W>, 0xf, 0xf, true);
  const double r = __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
  if constexpr (W == 32) return k == 0 ? 0.0 : r;
  return r;
}

// v moved one lane down by the W-lane DPP control (row_shr for 8 and 16,
// wave_shr:1 for 32 and 64, which crosses from lane 31 to lane 32); a lane
// without a source lane receives 0 (bound) or 'keep'.
template <int W>
__device__ __forceinline__ double dpp_down(double v, double keep, bool bound) {
  const long long vb = __double_as_longlong(v), kb = __double_as_longlong(keep);
  int lo, hi;
  if (bound) {
    lo = __builtin_amdgcn_update_dpp(0, (int)(vb & 0xffffffffLL), kDownCtrl<W>, 0xf, 0xf, true);
    hi = __builtin_amdgcn_update_dpp(0, (int)(vb >> 32), kDownCtrl<W>, 0xf, 0xf, true);
  } else {
    lo = __builtin_amdgcn_update_dpp((int)(kb & 0xffffffffLL), (int)(vb & 0xffffffffLL), kDownCtrl<W>, 0xf, 0xf,
                                     false);
    hi = __builtin_amdgcn_update_dpp((int)(kb >> 32), (int)(vb >> 32), kDownCtrl<W>, 0xf, 0xf, false);
  }
  return __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
}

template <typename T>
__device__ __forceinline__ const T* ph2pr_tab(const DevTables& t);
template <>
__device__ __forceinline__ const float* ph2pr_tab<float>(const DevTables& t) { return t.ph2pr_f; }
template <>
__device__ __forceinline__ const double* ph2pr_tab<double>(const DevTables& t) { return t.ph2pr_d; }
template <typename T>
__device__ __forceinline__ const T* mm_tab(const DevTables& t);
template <>
__device__ __forceinline__ const float* mm_tab<float>(const DevTables& t) { return t.mm_f; }
template <>
__device__ __forceinline__ const double* mm_tab<double>(const DevTables& t) { return t.mm_d; }

// Decode haplotype bases [0, H) into LDS codes: bits 0-1 base, bit 2 N.
__device__ __forceinline__ void stage_hap(const DevSeqs& s, uint32_t hap, uint32_t H, uint8_t* dst, int first,
                                          int step) {
  const uint32_t off = s.hap_off[hap];
  for (uint32_t x = first; x < H; x += step) {
    const uint32_t g = off + x;
    const uint32_t b = (s.hap_pack[g >> 4] >> ((g & 15u) * 2u)) & 3u;
    const uint32_t n = (s.hap_nmask[g >> 5] >> (g & 31u)) & 1u;
    dst[x] = (uint8_t)(b | (n << 2));
  }
}

// I = tMI*M_up + tII*I_up in the plain form. In float the product off the
// row-to-row chain (tMI*M_up) is rounded and tII*I_up is fused, so the chain
// down a lane's rows is one FMA per row; in double the other way round keeps
// the plain FP64 kernels within 256 registers (two waves per SIMD).
template <typename T>
__device__ __forceinline__ T i_update(T tmi, T um, T tii, T ui);
template <>
__device__ __forceinline__ float i_update<float>(float tmi, float um, float tii, float ui) {
  return fmaf(tii, ui, tmi * um);
}
template <>
__device__ __forceinline__ double i_update<double>(double tmi, double um, double tii, double ui) {
  return fma(tmi, um, tii * ui);
}

// Forms of the recurrence a sweep runs (template argument F).
//  kPlain  the model's M, I and D, 8 operations per cell:
//            M = prior * (tMM*M' + tGM*(I' + D'))
//            I = tMI*M_up + tII*I_up      D = tMD*M_left + tDD*D_left
//  kScaled I and D stored divided by their own row's gap-open probability,
//          I~ = I / tMI[r] and D~ = D / tMD[r] (row -1: D~ = D), M unchanged:
//            M  = (prior*tMM) * (M' + rho*I~' + delta*D~')
//            I~ = M_up + c*I~_up         D~ = M_left + tDD*D~_left
//          with rho = tGM*tMI[r-1]/tMM, delta = tGM*tMD[r-1]/tMM and
//          c = tII*tMI[r-1]/tMI[r]: 5 operations per cell. In this form the
//          coefficient arrays hold tmi = rho, tmd = delta, tii = c, tdd = tDD
//          and the priors are multiplied by tMM (row_params). I~ exceeds M by
//          up to the ratio of the largest to the smallest gap-open
//          probability in the region, so the host lowers the row -1 scale
//          2^120 by that ratio to 2^X (X per region, carried by each tile and
//          rescue entry) and the sum is scaled back by 2^(120 - X), exactly,
//          before the rescue test. Needs tMM > 0 and tMI, tMD > 0,
//          which the host's range check on the gap qualities ensures.
constexpr int kPlain = 0, kScaled = 1;

// 2^e, exactly, for e in the normal range of the type.
__device__ __forceinline__ float pow2f(int e) { return __builtin_bit_cast(float, (uint32_t)(e + 127) << 23); }
__device__ __forceinline__ double pow2d(int e) {
  return __builtin_bit_cast(double, (uint64_t)(e + 1023) << 52);
}

// New M, I, D of one cell from the row above at column j-1 (d*), the row
// above at column j (u*) and the same row at column j-1 (l*).
template <typename T, int F>
__device__ __forceinline__ void cell(T pr, T dm, T di, T dd, T um, T ui, T lm, T ld, T tmm, T tgm, T tmi, T tii,
                                     T tmd, T tdd, T& nm, T& ni, T& nd) {
  if constexpr (F == kScaled) {
    nm = pr * fma(tmd, dd, fma(tmi, di, dm));
    ni = fma(tii, ui, um);
    nd = fma(tdd, ld, lm);
  } else {
    nm = pr * fma(tmm, dm, tgm * (di + dd));
    ni = i_update<T>(tmi, um, tii, ui);
    nd = fma(tmd, lm, tdd * ld);
  }
}

// One anti-diagonal step of a lane: new column j of all RPL rows from
// column j-1 (M, I, D) and the row above (d* at j-1, u* at j).
template <typename T, int RPL, int F>
__device__ __forceinline__ void sweep_step(const T (&M)[RPL], const T (&I)[RPL], const T (&D)[RPL], T (&nM)[RPL],
                                           T (&nI)[RPL], T (&nD)[RPL], T dM, T dI, T dD, T uM, T uI, int hc,
                                           const int (&rcode)[RPL], const T (&pm)[RPL], const T (&px)[RPL],
                                           const T (&tmm)[RPL], const T (&tgm)[RPL], const T (&tmi)[RPL],
                                           const T (&tii)[RPL], const T (&tmd)[RPL], const T (&tdd)[RPL]) {
  const bool hn = (hc & 4) != 0;
  {
    const bool mt = (hc == rcode[0]) || hn;
    const T pr = mt ? pm[0] : px[0];
    cell<T, F>(pr, dM, dI, dD, uM, uI, M[0], D[0], tmm[0], tgm[0], tmi[0], tii[0], tmd[0], tdd[0], nM[0], nI[0],
               nD[0]);
  }
#pragma unroll
  for (int r = 1; r < RPL; ++r) {
    const bool mt = (hc == rcode[r]) || hn;
    const T pr = mt ? pm[r] : px[r];
    cell<T, F>(pr, M[r - 1], I[r - 1], D[r - 1], nM[r - 1], nI[r - 1], M[r], D[r], tmm[r], tgm[r], tmi[r], tii[r],
               tmd[r], tdd[r], nM[r], nI[r], nD[r]);
  }
}

// Transition parameters of lane k's rows k*RPL + r (r < RPL) of a read, and
// per row (r < NR) the match and mismatch priors and the base code (-1 for
// N, which matches every haplotype base). Rows past the read are padding
// rows: prior 0, I carries M + I of the row above, D stays 0.
// (ph/roDV_inter/pairhmm_inter.cpp:60-79; the forms are described above.)
template <typename T, int RPL, int NR, int F>
__device__ __forceinline__ void row_params(const DevSeqs& s, const T* ph2pr, const T* mmt, uint32_t read, int k,
                                           T (&tmm)[RPL], T (&tgm)[RPL], T (&tmi)[RPL], T (&tii)[RPL],
                                           T (&tmd)[RPL], T (&tdd)[RPL], T (&pm)[NR], T (&px)[NR], int (&rc)[NR]) {
  const uint32_t R = s.read_len[read];
  const uint32_t roff = s.read_off[read];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const uint32_t row = (uint32_t)k * RPL + r;
    pm[r] = 0; px[r] = 0; rc[r] = 0;
    if (r >= RPL) continue;
    if (row < R) {
      const uint32_t g = roff + row;
      const int iq = s.rq[g] & 127, ii = s.ri[g] & 127, id = s.rd[g] & 127, ic = s.rc[g] & 127;
      const uint32_t base = (s.read_pack[g >> 4] >> ((g & 15u) * 2u)) & 3u;
      const uint32_t isn = (s.read_nmask[g >> 5] >> (g & 31u)) & 1u;
      const T dist = ph2pr[iq];
      pm[r] = (T)1 - dist;
      px[r] = isn ? pm[r] : dist * ((T)1 / (T)3);
      rc[r] = isn ? -1 : (int)base;
      const T gc = ph2pr[ic];
      const int mx = ii > id ? ii : id, mn = ii > id ? id : ii;
      tmm[r] = mmt[((mx * (mx + 1)) >> 1) + mn];
      tgm[r] = (T)1 - gc;
      if constexpr (F == kScaled) {
        // row -1 has I = 0 and D~ = D
        const T aprev = row > 0 ? ph2pr[s.ri[g - 1] & 127] : (T)1;
        const T bprev = row > 0 ? ph2pr[s.rd[g - 1] & 127] : (T)1;
        pm[r] *= tmm[r];
        px[r] *= tmm[r];
        tmi[r] = tgm[r] * aprev / tmm[r];
        tmd[r] = tgm[r] * bprev / tmm[r];
        tii[r] = gc * aprev / ph2pr[ii];
        tdd[r] = gc;
      } else {
        tmi[r] = ph2pr[ii];
        tii[r] = gc;
        tmd[r] = ph2pr[id];
        tdd[r] = gc;
      }
    } else {
      tmm[r] = 0; tgm[r] = 0; tmi[r] = 1; tii[r] = 1; tmd[r] = 0; tdd[r] = 1;
      if constexpr (F == kScaled) {
        // the first padding row holds M + I of the read's last row: its I~
        // (scale 1) takes that row's I~ times the row's tMI
        tmi[r] = 0;
        tii[r] = row == R ? ph2pr[s.ri[roff + R - 1] & 127] : (T)1;
      }
    }
  }
}

// S sums M + f*I of the segment's last row: f = 1, or in the scaled form
// tMI of that row when it is the read's last row (R = W * RPL, no padding
// row; the FP64 bins allow that), since there I holds I~ = I / tMI.
template <typename T, int W, int RPL, int F>
__device__ __forceinline__ T sum_factor(const DevSeqs& s, const T* ph2pr, uint32_t read) {
  const uint32_t R = s.read_len[read];
  if (F != kScaled || R < (uint32_t)(W * RPL)) return (T)1;
  return ph2pr[s.ri[s.read_off[read] + R - 1] & 127];
}

// Forward sweep of one pair per segment. hb points at the segment's staged
// haplotype codes (hb[j] = code of base j). Returns the raw sum S in the
// segment's last lane; H is this segment's haplotype length, Hmax the
// largest among the wave's segments. c0 is the row -1 scale (2^120 in float,
// 2^1020 in double, lower in the scaled form).
template <typename T, int W, int RPL, int F>
__device__ __forceinline__ T forward(const DevSeqs& s, const DevTables& tb, const uint8_t* hb, int k, uint32_t H,
                                     uint32_t Hmax, uint32_t read, T c0) {
  const T* ph2pr = ph2pr_tab<T>(tb);
  const T* mmt = mm_tab<T>(tb);

  T pm[RPL], px[RPL], tmm[RPL], tgm[RPL], tmi[RPL], tii[RPL], tmd[RPL], tdd[RPL];
  int rcode[RPL];
  row_params<T, RPL, RPL, F>(s, ph2pr, mmt, read, k, tmm, tgm, tmi, tii, tmd, tdd, pm, px, rcode);

  // Row -1 deletion value c0/H, fed to lane 0.
  const T initD = c0 / (T)H;
  const T old_d = initD;
  const T sfac = sum_factor<T, W, RPL, F>(s, ph2pr, read);
  T Ma[RPL], Ia[RPL], Da[RPL], Mb[RPL], Ib[RPL], Db[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) { Ma[r] = 0; Ia[r] = 0; Da[r] = 0; }
  // Row above this lane's first row: at column j (u*) and column j-1 (d*).
  T uM = 0, uI = 0, uD = (k == 0) ? initD : (T)0;
  T dM = 0, dI = 0, dD = uD;
  T S = 0;
  // Two steps per iteration with ping-pong registers (a -> b -> a), so the
  // loop carries no register copies. A step past the last column only
  // touches columns >= H, which nothing reads, so the count is rounded up.
  const int nsteps = (int)Hmax + W - 1;
  const int mysteps = (int)H + W - 1;
  const uint8_t* hp = hb - k;  // hp[s] = code of column s - k
  for (int st = 0; st < nsteps; st += 2) {
    const int hc0 = hp[st];
    const int hc1 = hp[st + 1];
    sweep_step<T, RPL, F>(Ma, Ia, Da, Mb, Ib, Db, dM, dI, dD, uM, uI, hc0, rcode, pm, px, tmm, tgm, tmi, tii, tmd,
                          tdd);
    if (st < mysteps) S += fma(sfac, Ib[RPL - 1], Mb[RPL - 1]);
    const T sM = lane_down0<W>(Mb[RPL - 1], k);
    const T sI = lane_down0<W>(Ib[RPL - 1], k);
    const T sD = lane_down<W>(Db[RPL - 1], old_d, k);
    sweep_step<T, RPL, F>(Mb, Ib, Db, Ma, Ia, Da, uM, uI, uD, sM, sI, hc1, rcode, pm, px, tmm, tgm, tmi, tii, tmd,
                          tdd);
    if (st + 1 < mysteps) S += fma(sfac, Ia[RPL - 1], Ma[RPL - 1]);
    dM = sM; dI = sI; dD = sD;
    uM = lane_down0<W>(Ma[RPL - 1], k);
    uI = lane_down0<W>(Ia[RPL - 1], k);
    uD = lane_down<W>(Da[RPL - 1], old_d, k);
  }
  return S;
}

// Result of one FP32 pair (output slot o): log10 likelihood, or a place on
// the FP64 list.
// x is the pair's scale exponent (S already brought back to the 2^120 scale).
__device__ __forceinline__ bool store_fp32(const Fp32Args& a, uint32_t o, float S) {
  if (S >= 1e-28f && S <= 0x1.fffffep127f) {
    a.out[o] = log10((double)S) - (double)36.1236000061f;
    a.flags[o] = 1;
    return true;
  }
  return false;
}
__device__ __forceinline__ void finish_fp32(const Fp32Args& a, uint32_t o, uint32_t hap, uint32_t read, uint32_t x,
                                            float S) {
  if (!store_fp32(a, o, S)) {
    const int b = (int)a.rescue_list0 + fp64_bin(a.s.read_len[read]);
    const uint32_t idx = atomicAdd(a.rescue_count + b, 1u);
    RescueEntry e;
    e.hap = hap; e.read = read; e.out = o; e.x = x;
    a.rescue[a.rescue_base[b] + idx] = e;
  }
}

// FP32 pass: persistent waves pull tiles (a run of haplotypes x up to 64/W
// reads). OCC = waves per SIMD the register allocation must allow (512 / OCC
// VGPRs per lane): the sweep is issue-latency bound, so occupancy is a tuning
// knob. XF selects the scaled form, else the plain one.
template <int W, int RPL, bool XF, int OCC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC))) void k_fp32(Fp32Args a) {
  constexpr int F = XF ? kScaled : kPlain;
  __shared__ uint8_t hbuf[kHapBuf];
  const int lane = threadIdx.x;
  const int seg = seg_of<W>(lane), k = pos_of<W>(lane);
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.head, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.n_tiles) break;
    const Tile tl = a.tiles[t];
    const bool active = seg < (int)tile_nseg(tl.meta);
    const uint32_t stride = tile_stride(tl.meta), nh = tile_nh(tl.meta);
    const uint32_t read = tl.read0 + (active ? (uint32_t)seg : 0u);
    const uint32_t X = tile_x(tl.meta);
    const float init = pow2f((int)X), back = pow2f(120 - (int)X);
    for (uint32_t h = 0; h < nh; ++h) {
      const uint32_t hap = tl.hap + h;
      const uint32_t H = a.s.hap_len[hap];
      __syncthreads();
      stage_hap(a.s, hap, H, hbuf + 64, lane, 64);
      __syncthreads();
      const float S = forward<float, W, RPL, F>(a.s, a.t, hbuf + 64, k, H, H, read, init);
      if (active && k == W - 1) finish_fp32(a, tl.out0 + (uint32_t)seg * stride + h, hap, read, X, S * back);
    }
  }
}

// ---------------------------------------------------------------------------
// Prior-table variant of the FP32 pass, for haplotypes without N.
//
// The prior of a cell depends only on its read row and on the haplotype base
// (4 values), so each tile first writes, per lane and row, the prior for each
// of the 4 bases into LDS as ptab[base][row/4][lane] (float4 = 4 rows). A step
// then reads the 4-row groups of its column's base with one ds_read_b128 each,
// one step ahead, instead of a compare and a select per cell. The haplotype
// itself sits in LDS as its 2-bit stream (16 bases per word, 64 zero columns
// in front); a funnel shift yields the codes of 4 consecutive columns.
// ---------------------------------------------------------------------------
constexpr int kHapWords = 112;  // 64 pad columns + 1536 bases + tail, 16 per word
// Most consecutive rescue-list entries one segment of the FP64 prior-table
// kernel takes per fetch (fewer on short lists).
constexpr uint32_t kFp64Chunk = 8;

template <typename T, int RPL, int F>
__device__ __forceinline__ void sweep_step_pt(const T (&M)[RPL], const T (&I)[RPL], const T (&D)[RPL], T (&nM)[RPL],
                                              T (&nI)[RPL], T (&nD)[RPL], T dM, T dI, T dD, T uM, T uI, const T* pr,
                                              const T (&tmm)[RPL], const T (&tgm)[RPL], const T (&tmi)[RPL],
                                              const T (&tii)[RPL], const T (&tmd)[RPL], const T (&tdd)[RPL]) {
  cell<T, F>(pr[0], dM, dI, dD, uM, uI, M[0], D[0], tmm[0], tgm[0], tmi[0], tii[0], tmd[0], tdd[0], nM[0], nI[0],
             nD[0]);
#pragma unroll
  for (int r = 1; r < RPL; ++r)
    cell<T, F>(pr[r], M[r - 1], I[r - 1], D[r - 1], nM[r - 1], nI[r - 1], M[r], D[r], tmm[r], tgm[r], tmi[r], tii[r],
               tmd[r], tdd[r], nM[r], nI[r], nD[r]);
}

// One step of the prior-table sweep where row 0 of a lane takes its inputs
// premixed by the lane above, with this lane's row-0 coefficients:
//   Y = M' + rho*I~' + delta*D~'   (the lane above's last row, two steps old: the diagonal)
//   X = M_up + c*I~_up            (one step old: the row above in this column)
// so M[0] = prior * Y and I~[0] = X are exactly the operations of the full
// cell, done by the sender, and a step moves two values between lanes
// instead of three (M, I and D).
template <typename T, int RPL>
__device__ __forceinline__ void sweep_step_xy(const T (&M)[RPL], const T (&I)[RPL], const T (&D)[RPL], T (&nM)[RPL],
                                              T (&nI)[RPL], T (&nD)[RPL], T yd, T xu, const T* pr,
                                              const T (&tmm)[RPL], const T (&tgm)[RPL], const T (&tmi)[RPL],
                                              const T (&tii)[RPL], const T (&tmd)[RPL], const T (&tdd)[RPL]) {
  nM[0] = pr[0] * yd;
  nI[0] = xu;
  nD[0] = fma(tdd[0], D[0], M[0]);
#pragma unroll
  for (int r = 1; r < RPL; ++r)
    cell<T, kScaled>(pr[r], M[r - 1], I[r - 1], D[r - 1], nM[r - 1], nI[r - 1], M[r], D[r], tmm[r], tgm[r], tmi[r],
                     tii[r], tmd[r], tdd[r], nM[r], nI[r], nD[r]);
}

// v of lane (lane + delta) mod 64, through the LDS crossbar (ds_bpermute:
// no LDS memory is touched, every lane reads its source lane's register).
__device__ __forceinline__ float lane_from(float v, int delta) {
  const int addr = ((int)(threadIdx.x + delta) & 63) << 2;
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double lane_from(double v, int delta) {
  const int addr = ((int)(threadIdx.x + delta) & 63) << 2;
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
}

template <int W, int RPL, bool XF, int OCC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC))) void k_fp32pt(Fp32Args a) {
  constexpr int NQ = (RPL + 3) / 4;
  constexpr int F = XF ? kScaled : kPlain;
  // Past 16 rows per lane the last, partial group of 4 rows is stored as
  // single floats (TR of them): 19 rows then take 19.5 KB of table instead
  // of 20.5 KB, and eight waves fit in a CU's 160 KB of LDS.
  constexpr int TR = RPL > 16 ? RPL % 4 : 0;  // rows in the single-float tail
  constexpr int NF = TR ? RPL / 4 : NQ;       // float4 groups
  __shared__ float4 ptab[4 * NF * 64];
  __shared__ float ptail[TR ? 4 * TR * 64 : 1];
  __shared__ uint32_t hpk[kHapWords];
  const int lane = threadIdx.x;
  const int seg = seg_of<W>(lane), k = pos_of<W>(lane);
  const float* ph2pr = a.t.ph2pr_f;
  const float* mmt = a.t.mm_f;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.head, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.n_tiles) break;
    const Tile tl = a.tiles[t];
    const bool active = seg < (int)tile_nseg(tl.meta);
    const uint32_t stride = tile_stride(tl.meta), nh = tile_nh(tl.meta), X = tile_x(tl.meta);
    const uint32_t read = tl.read0 + (active ? (uint32_t)seg : 0u);
    __syncthreads();
    // per-row transition parameters (registers) and priors (LDS table), once
    // for all haplotypes of the tile
    float tmm[RPL], tgm[RPL], tmi[RPL], tii[RPL], tmd[RPL], tdd[RPL];
    float pm[NQ * 4], px[NQ * 4];
    int rc[NQ * 4];
    row_params<float, RPL, NQ * 4, F>(a.s, ph2pr, mmt, read, k, tmm, tgm, tmi, tii, tmd, tdd, pm, px, rc);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
      for (int q = 0; q < NF; ++q) {
        float4 v;
        v.x = (rc[4 * q + 0] < 0 || rc[4 * q + 0] == c) ? pm[4 * q + 0] : px[4 * q + 0];
        v.y = (rc[4 * q + 1] < 0 || rc[4 * q + 1] == c) ? pm[4 * q + 1] : px[4 * q + 1];
        v.z = (rc[4 * q + 2] < 0 || rc[4 * q + 2] == c) ? pm[4 * q + 2] : px[4 * q + 2];
        v.w = (rc[4 * q + 3] < 0 || rc[4 * q + 3] == c) ? pm[4 * q + 3] : px[4 * q + 3];
        ptab[(c * NF + q) * 64 + lane] = v;
      }
#pragma unroll
      for (int t = 0; t < TR; ++t) {
        const int r = 4 * NF + t;
        ptail[(c * TR + t) * 64 + lane] = (rc[r] < 0 || rc[r] == c) ? pm[r] : px[r];
      }
    }
    // Row-0 coefficients of the lane below, which receives this lane's X and
    // Y. A segment's last lane feeds the next segment's first lane (or wraps
    // to lane 0), which must see row -1 (M = I = 0, D~ = initD): its last row
    // is a padding row (M = 0, D~ held at initD from the start), so with
    // c = rho = 0 it sends X = 0 and Y = delta * initD. (The lane permutes
    // run in every lane: each reads its source lane's register.)
    constexpr int SH = W == 8 ? 2 : 1;  // lane distance to the next position of a segment
    const float cB = lane_from(tii[0], SH), rB = lane_from(tmi[0], SH);
    const float dN = lane_from(tmd[0], SH);
    const float cN = k == W - 1 ? 0.f : cB;
    const float rN = k == W - 1 ? 0.f : rB;
    // this lane's words x = lane, lane + 64 of a haplotype's stream (64
    // zero columns in front); the next haplotype's are loaded while the
    // current one is swept
    auto hap_words = [&](uint32_t hap, uint32_t (&w)[2]) {
      const uint32_t Hh = a.s.hap_len[hap];
      const uint32_t w0 = a.s.hap_off[hap] >> 4, nw = (Hh + 15) >> 4;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int x = lane + 64 * i, hx = x - 4;
        w[i] = (x < kHapWords && hx >= 0 && hx < (int)nw) ? a.s.hap_pack[w0 + hx] : 0u;
      }
    };
    uint32_t words[2];
    hap_words(tl.hap, words);
    static_assert(kMaxHapRun <= 64, "a run's rescue mask has one bit per haplotype");
    uint64_t resc = 0;  // the run's haplotypes whose pair goes to the FP64 rescue
    for (uint32_t h = 0; h < nh; ++h) {
      const uint32_t hap = tl.hap + h;
      const uint32_t H = a.s.hap_len[hap];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 2; ++i)
        if (lane + 64 * i < kHapWords) hpk[lane + 64 * i] = words[i];
      __syncthreads();
      if (h + 1 < nh) hap_words(hap + 1, words);

      const float initD = pow2f((int)X) / (float)H;
      float Ma[RPL], Ia[RPL], Da[RPL], Mb[RPL], Ib[RPL], Db[RPL];
#pragma unroll
      for (int r = 0; r < RPL; ++r) { Ma[r] = 0.f; Ia[r] = 0.f; Da[r] = 0.f; }
      Da[RPL - 1] = k == W - 1 ? initD : 0.f;
      // Y of the two steps before step 0 and X of the step before: row -1
      // for a segment's first lane, zeros elsewhere (columns < 0)
      const float y0 = k == 0 ? tmd[0] * initD : 0.f;
      float yd = y0, yn = y0, xu = 0.f;
      float S = 0.f;
      const int nsteps = (int)H + W - 1;
      const char* tab = reinterpret_cast<const char*>(ptab) + lane * 16;
      // codes of columns st-k .. st-k+3 (stream position p = column + 64 >= 1)
      auto window = [&](int st) -> uint32_t {
        const int p = st - k + 64;
        return __builtin_amdgcn_alignbit(hpk[(p >> 4) + 1], hpk[p >> 4], (uint32_t)(p & 15) * 2u);
      };
      auto load = [&](uint32_t codes, int i, float (&pr)[NQ * 4]) {
        const uint32_t c = (codes >> (2 * i)) & 3u;
        const float4* b = reinterpret_cast<const float4*>(tab + c * (uint32_t)(NF * 1024));
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const float4 v = b[q * 64];
          pr[4 * q + 0] = v.x; pr[4 * q + 1] = v.y; pr[4 * q + 2] = v.z; pr[4 * q + 3] = v.w;
        }
#pragma unroll
        for (int t = 0; t < TR; ++t) pr[4 * NF + t] = ptail[(c * TR + t) * 64 + lane];
      };
      // After a step: this lane's X and Y for the lane below, moved there.
      // 16-lane segments move them with LDS lane permutes, 8-lane segments
      // (and 32/64) with DPP, where a segment's first lane takes X = 0 (bound
      // control) and keeps Y = y0, which the register of two steps before
      // (dead after this step) already holds there (measured on the batched
      // workload: 16 x 16 49.2 -> 46.7 ms with permutes, 47.6 with DPP;
      // 8 x 19 36.1 -> 34.3 ms with DPP, 36.0 with permutes).
      auto send = [&](const float (&nM)[RPL], const float (&nI)[RPL], const float (&nD)[RPL]) {
        const float x = fmaf(cN, nI[RPL - 1], nM[RPL - 1]);
        const float y = fmaf(dN, nD[RPL - 1], fmaf(rN, nI[RPL - 1], nM[RPL - 1]));
        if constexpr (W == 16) {
          xu = lane_from(x, -SH);
          yd = yn;
          yn = lane_from(y, -SH);
        } else {
          xu = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kDownCtrl<W>, 0xf, 0xf, true));
          const float t = __int_as_float(
              __builtin_amdgcn_update_dpp(__float_as_int(yd), __float_as_int(y), kDownCtrl<W>, 0xf, 0xf, false));
          yd = yn;
          yn = t;
        }
      };
      // Four steps a -> b -> a -> b -> a per iteration. The read's last row is
      // always a padding row (the host picks RPL with R < W * RPL), which holds
      // M[R-1][j] + I[R-1][j] in I, so the sum takes one add per step.
      // sched_barrier(0) keeps each step's table reads ahead of the previous
      // step's arithmetic, so their latency overlaps a whole step.
      // Short steps (up to 8 rows per lane) do not cover a table read, and at
      // two waves per SIMD (OCC 2) up to 13 rows leave registers for more
      // buffers: those read three steps ahead (measured: 8 lanes x 13 rows
      // 22.4 -> 21.8 ms on the batched workload; 16 x 10 at OCC 3 loses).
      if constexpr (NQ > 2 && !(OCC == 2 && RPL <= 13)) {
        float pA[NQ * 4], pB[NQ * 4];
        uint32_t cw = window(0);
        load(cw, 0, pA);
        for (int st = 0; st < nsteps; st += 4) {
          const uint32_t cw_next = window(st + 4);
          load(cw, 1, pB);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Ma, Ia, Da, Mb, Ib, Db, yd, xu, pA, tmm, tgm, tmi, tii, tmd, tdd);
          S += Ib[RPL - 1];
          send(Mb, Ib, Db);
          load(cw, 2, pA);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Mb, Ib, Db, Ma, Ia, Da, yd, xu, pB, tmm, tgm, tmi, tii, tmd, tdd);
          if (st + 1 < nsteps) S += Ia[RPL - 1];
          send(Ma, Ia, Da);
          load(cw, 3, pB);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Ma, Ia, Da, Mb, Ib, Db, yd, xu, pA, tmm, tgm, tmi, tii, tmd, tdd);
          if (st + 2 < nsteps) S += Ib[RPL - 1];
          send(Mb, Ib, Db);
          cw = cw_next;
          load(cw, 0, pA);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Mb, Ib, Db, Ma, Ia, Da, yd, xu, pB, tmm, tgm, tmi, tii, tmd, tdd);
          if (st + 3 < nsteps) S += Ia[RPL - 1];
          send(Ma, Ia, Da);
        }
      } else {
        // three steps ahead, four rotating buffers
        float p0[NQ * 4], p1[NQ * 4], p2[NQ * 4], p3[NQ * 4];
        uint32_t cw = window(0), cw1 = window(4);
        load(cw, 0, p0);
        load(cw, 1, p1);
        load(cw, 2, p2);
        for (int st = 0; st < nsteps; st += 4) {
          const uint32_t cw2 = window(st + 8);
          load(cw, 3, p3);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Ma, Ia, Da, Mb, Ib, Db, yd, xu, p0, tmm, tgm, tmi, tii, tmd, tdd);
          S += Ib[RPL - 1];
          send(Mb, Ib, Db);
          load(cw1, 0, p0);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Mb, Ib, Db, Ma, Ia, Da, yd, xu, p1, tmm, tgm, tmi, tii, tmd, tdd);
          if (st + 1 < nsteps) S += Ia[RPL - 1];
          send(Ma, Ia, Da);
          load(cw1, 1, p1);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Ma, Ia, Da, Mb, Ib, Db, yd, xu, p2, tmm, tgm, tmi, tii, tmd, tdd);
          if (st + 2 < nsteps) S += Ib[RPL - 1];
          send(Mb, Ib, Db);
          load(cw1, 2, p2);
          __builtin_amdgcn_sched_barrier(0);
          sweep_step_xy<float, RPL>(Mb, Ib, Db, Ma, Ia, Da, yd, xu, p3, tmm, tgm, tmi, tii, tmd, tdd);
          if (st + 3 < nsteps) S += Ia[RPL - 1];
          send(Ma, Ia, Da);
          cw = cw1;
          cw1 = cw2;
        }
      }
      if (active && k == W - 1 && !store_fp32(a, tl.out0 + (uint32_t)seg * stride + h, S * pow2f(120 - (int)X)))
        resc |= 1ull << h;
    }
    // the run's pairs below the FP32 range go onto the rescue list together,
    // one read's haplotypes next to each other, so that the FP64 pass sets
    // each read up once for several of its pairs
    if (active && k == W - 1 && resc) {
      const int b = (int)a.rescue_list0 + fp64_bin(a.s.read_len[read]);
      uint32_t idx = a.rescue_base[b] + atomicAdd(a.rescue_count + b, (uint32_t)__popcll(resc));
      for (uint64_t m = resc; m; m &= m - 1) {
        const uint32_t h = (uint32_t)__ffsll((long long)m) - 1;
        RescueEntry e;
        e.hap = tl.hap + h; e.read = read; e.out = tl.out0 + (uint32_t)seg * stride + h; e.x = X;
        a.rescue[idx++] = e;
      }
    }
  }
}

// FP64 rescue with the prior table, for pairs whose haplotype has no N, in
// the scaled form: the prior-table sweep in double (2 rows per ds_read_b128).
// The segments of a wave hold different pairs, so each segment stages its own
// haplotype stream and the wave sweeps to the longest of them; a segment adds
// to its sum only while its own columns last (the test is left out of the
// loop while every segment is inside its own haplotype).
template <int W, int RPL>
__global__ __launch_bounds__(64) void k_fp64pt(Fp64Args a) {
  constexpr int NSEG = 64 / W;
  constexpr int NP = (RPL + 1) / 2;
  __shared__ double2 ptab[4 * NP * 64];
  __shared__ uint32_t hpk[NSEG][kHapWords];
  const int lane = threadIdx.x;
  const int seg = lane / W, k = lane % W;
  const double* ph2pr = a.t.ph2pr_d;
  const double* mmt = a.t.mm_d;
  const uint32_t count = *a.count;
  // Entries per segment and fetch. A read's pairs lie next to each other on
  // the list (the FP32 prior-table kernel appends a tile's pairs of one read
  // together), so a segment that takes several consecutive entries sets its
  // read up (row parameters, prior table, next-lane coefficients) once for
  // all of them; a short list keeps one entry per segment, so that every
  // wave has work.
  const uint32_t chunk = min(kFp64Chunk, max(1u, count / (gridDim.x * (uint32_t)NSEG * 4u)));
  constexpr bool kPad = W < 64;
  double tmm[RPL], tgm[RPL], tmi[RPL], tii[RPL], tmd[RPL], tdd[RPL];
  double sfac = 1.0, cN = 0.0, rN = 0.0, dN = 0.0;
  uint32_t cur_read = 0xffffffffu;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.head, (uint32_t)NSEG * chunk);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= count) break;
    for (uint32_t j = 0; j < chunk; ++j) {
      if (t + j >= count) break;  // segment 0 holds the lowest entries of the fetch
      const uint32_t e = t + (uint32_t)seg * chunk + j;
      const bool active = e < count;
      const RescueEntry en = a.list[active ? e : t + j];
      const uint32_t H = a.s.hap_len[en.hap];
      uint32_t Hmax = H, Hmin = H;
#pragma unroll
      for (int g = 0; g < NSEG; ++g) {
        const uint32_t h2 = (uint32_t)__shfl((int)H, g * W);
        Hmax = h2 > Hmax ? h2 : Hmax;
        Hmin = h2 < Hmin ? h2 : Hmin;
      }
      __syncthreads();
      {
        const uint32_t w0 = a.s.hap_off[en.hap] >> 4, nw = (H + 15) >> 4;
        for (int x = k; x < kHapWords; x += W) {
          const int hx = x - 4;
          hpk[seg][x] = (hx >= 0 && hx < (int)nw) ? a.s.hap_pack[w0 + hx] : 0u;
        }
      }
      // a new read in some segment: the whole wave sets its reads up again
      // (a segment whose read is unchanged computes the same values)
      if (__ballot(en.read != cur_read) != 0) {
        cur_read = en.read;
        double pm[NP * 2], px[NP * 2];
        int rc[NP * 2];
        row_params<double, RPL, NP * 2, kScaled>(a.s, ph2pr, mmt, en.read, k, tmm, tgm, tmi, tii, tmd, tdd, pm, px,
                                                 rc);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int q = 0; q < NP; ++q) {
            double2 v;
            v.x = (rc[2 * q + 0] < 0 || rc[2 * q + 0] == c) ? pm[2 * q + 0] : px[2 * q + 0];
            v.y = (rc[2 * q + 1] < 0 || rc[2 * q + 1] == c) ? pm[2 * q + 1] : px[2 * q + 1];
            ptab[(c * NP + q) * 64 + lane] = v;
          }
        // The 16- and 32-lane bins hold reads of fewer than W * RPL bases, so
        // a segment's last row is a padding row: M = 0, I~ carries M + I of
        // the read's last row (its sum term), D~ holds what it starts with.
        // The 64-lane bin takes reads of up to W * RPL bases and sums
        // M + f * I.
        sfac = kPad ? 1.0 : sum_factor<double, W, RPL, kScaled>(a.s, ph2pr, en.read);
        // Row 0 takes X and Y premixed by the lane above with this lane's
        // row-0 coefficients (as in the FP32 kernel): cN, rN, dN are the next
        // lane's.
        cN = lane_from(tii[0], 1);
        rN = lane_from(tmi[0], 1);
        dN = lane_from(tmd[0], 1);
        if (kPad && k == W - 1) { cN = 0.0; rN = 0.0; }
      }
      __syncthreads();

      const double initD = pow2d((int)en.x + 900) / (double)H;
      // A segment's first lane sees row -1 (X = 0, Y = delta * initD of its
      // own pair). DPP bound control gives lane 0 of the move X = 0 and the
      // kept register Y (16-lane rows, the 64-lane wave). In the 32-lane
      // wave-wide move lane 32 receives lane 31's values instead: lane 31's
      // padding row, with M = 0, next-lane c = rho = 0 and D~ held at the
      // second pair's initD, sends exactly X = 0 and Y = delta * initD.
      const double initDn = lane_from(initD, 1);
      double Ma[RPL], Ia[RPL], Da[RPL], Mb[RPL], Ib[RPL], Db[RPL];
#pragma unroll
      for (int r = 0; r < RPL; ++r) { Ma[r] = 0; Ia[r] = 0; Da[r] = 0; }
      if (kPad && k == W - 1) Da[RPL - 1] = initDn;
      const double y0 = k == 0 ? tmd[0] * initD : 0.0;
      double yd = y0, yn = y0, xu = 0.0;
      double S = 0;
      const int nsteps = (int)Hmax + W - 1;
      const int mysteps = (int)H + W - 1;
      const char* tab = reinterpret_cast<const char*>(ptab) + lane * 16;
      const uint32_t* hs = hpk[seg];
      auto codes_at = [&](int st) -> uint32_t {
        const int p = st - k + 64;
        return __builtin_amdgcn_alignbit(hs[(p >> 4) + 1], hs[p >> 4], (uint32_t)(p & 15) * 2u);
      };
      auto load = [&](uint32_t codes, int i, double (&pr)[NP * 2]) {
        const uint32_t c = (codes >> (2 * i)) & 3u;
        const double2* b = reinterpret_cast<const double2*>(tab + c * (uint32_t)(NP * 1024));
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          const double2 v = b[q * 64];
          pr[2 * q + 0] = v.x; pr[2 * q + 1] = v.y;
        }
      };
      auto send = [&](const double (&nM)[RPL], const double (&nI)[RPL], const double (&nD)[RPL]) {
        const double x = fma(cN, nI[RPL - 1], nM[RPL - 1]);
        const double y = fma(dN, nD[RPL - 1], fma(rN, nI[RPL - 1], nM[RPL - 1]));
        xu = dpp_down<W>(x, 0.0, true);
        const double t = dpp_down<W>(y, yd, false);  // yd (two steps old) dies here
        yd = yn;
        yn = t;
      };
      // the sum term of the last row: I~ of the padding row, or M + f * I
      auto term = [&](const double (&nM)[RPL], const double (&nI)[RPL]) -> double {
        return kPad ? nI[RPL - 1] : fma(sfac, nI[RPL - 1], nM[RPL - 1]);
      };
      double pA[NP * 2], pB[NP * 2];
      uint32_t cw = codes_at(0);
      load(cw, 0, pA);
      // four steps a -> b -> a -> b -> a per iteration, the table reads one
      // step ahead, the next codes word read once per iteration; while every
      // segment of the wave is inside its own columns the sums need no test
      auto iter4 = [&](int st, auto guarded) {
        constexpr bool G = decltype(guarded)::value;
        const uint32_t cw_next = codes_at(st + 4);
        load(cw, 1, pB);
        __builtin_amdgcn_sched_barrier(0);
        sweep_step_xy<double, RPL>(Ma, Ia, Da, Mb, Ib, Db, yd, xu, pA, tmm, tgm, tmi, tii, tmd, tdd);
        if (!G || st < mysteps) S += term(Mb, Ib);
        send(Mb, Ib, Db);
        load(cw, 2, pA);
        __builtin_amdgcn_sched_barrier(0);
        sweep_step_xy<double, RPL>(Mb, Ib, Db, Ma, Ia, Da, yd, xu, pB, tmm, tgm, tmi, tii, tmd, tdd);
        if (!G || st + 1 < mysteps) S += term(Ma, Ia);
        send(Ma, Ia, Da);
        load(cw, 3, pB);
        __builtin_amdgcn_sched_barrier(0);
        sweep_step_xy<double, RPL>(Ma, Ia, Da, Mb, Ib, Db, yd, xu, pA, tmm, tgm, tmi, tii, tmd, tdd);
        if (!G || st + 2 < mysteps) S += term(Mb, Ib);
        send(Mb, Ib, Db);
        cw = cw_next;
        load(cw, 0, pA);
        __builtin_amdgcn_sched_barrier(0);
        sweep_step_xy<double, RPL>(Mb, Ib, Db, Ma, Ia, Da, yd, xu, pB, tmm, tgm, tmi, tii, tmd, tdd);
        if (!G || st + 3 < mysteps) S += term(Ma, Ia);
        send(Ma, Ia, Da);
      };
      const int minsteps = (int)Hmin + W - 1;
      int st = 0;
      for (; st + 4 <= minsteps; st += 4) iter4(st, std::false_type{});
      for (; st < nsteps; st += 4) iter4(st, std::true_type{});
      if (active && k == W - 1) {
        a.out[en.out] = log10(S * pow2d(120 - (int)en.x)) - 307.050595577260822;
        a.flags[en.out] = 3;
      }
    }
  }
}

// FP64 rescue: persistent waves pull 64/W queued pairs at a time. SC selects
// the scaled form, else the plain one.
template <int W, int RPL, bool SC>
__global__ __launch_bounds__(64) void k_fp64(Fp64Args a) {
  constexpr int NSEG = 64 / W;
  __shared__ uint8_t hbuf[NSEG][kHapBuf];
  const int lane = threadIdx.x;
  const int seg = lane / W, k = lane % W;
  const uint32_t count = *a.count;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.head, (uint32_t)NSEG);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= count) break;
    const uint32_t e = t + seg;
    const bool active = e < count;
    const RescueEntry en = a.list[active ? e : t];
    const uint32_t H = a.s.hap_len[en.hap];
    uint32_t Hmax = H;
#pragma unroll
    for (int g = 0; g < NSEG; ++g) {
      const uint32_t h2 = (uint32_t)__shfl((int)H, g * W);
      Hmax = h2 > Hmax ? h2 : Hmax;
    }
    __syncthreads();
    stage_hap(a.s, en.hap, H, &hbuf[seg][64], k, W);
    __syncthreads();
    const double S = forward<double, W, RPL, SC ? kScaled : kPlain>(a.s, a.t, &hbuf[seg][64], k, H, Hmax, en.read,
                                                                    pow2d((int)en.x + 900));
    if (active && k == W - 1) {
      a.out[en.out] = log10(S * pow2d(120 - (int)en.x)) - 307.050595577260822;
      a.flags[en.out] = 3;
    }
  }
}

// Read preprocessing on the device, one wave per read (the read's bases are
// independent): q = min(qual, mq), 6 below min_q; gap open 45 lowered by the
// PCR indel model's cap for the base's tandem-repeat length (the last base
// keeps 45; insertions and deletions share it); gap continuation 10.
// ph/roDV_schedule.cpp:522-608, ph/pairhmm_internal.cpp:391-429,
// hc/genotype/read_record.cpp:184-210; the same steps as the host adapter.
__global__ __launch_bounds__(64) void k_prep(PrepArgs a) {
  __shared__ uint8_t bases[kMaxReadLen];
  const uint32_t r = blockIdx.x;
  const uint32_t w = a.prep[r];
  if (!(w & kPrepOn)) return;
  const int mq = (int)(w & 255u), min_q = (int)((w >> 8) & 255u), pcr = (int)((w >> 16) & 3u);
  const int R = (int)a.read_len[r];
  const uint32_t off = a.read_off[r];
  // the repeat search reads each base many times: from LDS, not from L2
  for (int x = threadIdx.x; x < R; x += 64) bases[x] = a.raw[off + x];
  __syncthreads();
  for (int pos = threadIdx.x; pos < R; pos += 64) {
    const uint32_t g = off + (uint32_t)pos;
    const int qual = a.rq[g];
    const int capped = qual < mq ? qual : mq;
    int gop = '-';
    if (pcr != 0 && pos + 1 < R) {
      const int cap = a.gop_cap[(pcr - 1) * 21 + repeat_length(bases, R, pos)];
      gop = cap < gop ? cap : gop;
    }
    a.rq[g] = (uint8_t)(capped < min_q ? 6 : capped);
    a.ri[g] = (uint8_t)gop;
    a.rd[g] = (uint8_t)gop;
    a.rc[g] = (uint8_t)('+' - 33);
  }
}

// One block per region, a thread per read: normalise the read's row as
// ph/pairhmm_internal.cpp:248-258 does (best found as std::max_element finds
// it, so a leading NaN leaves the row as it is; values below best - 4.5 are
// raised to it), then for every site and allele the largest of the values of
// the haplotypes mapped to it, starting from -inf, with a strict compare so
// that a NaN never wins (allele_likelihoods.hpp:713-729).
__global__ __launch_bounds__(256) void k_marginalize(MargArgs a) {
  const MargRegion g = a.regions[blockIdx.x];
  const uint32_t H = g.n_haps;
  for (uint32_t sx = threadIdx.x; sx < g.n_reads; sx += blockDim.x) {
    const double* row = a.results + g.pair0 + (uint64_t)sx * H;
    double best = row[0];
    for (uint32_t h = 1; h < H; ++h)
      if (best < row[h]) best = row[h];
    const double floor_v = best + (-4.5);
    const uint32_t x = a.read_perm[g.read0 + sx];
    for (uint32_t s = g.site0; s < g.site0 + g.n_sites; ++s) {
      const int16_t* m = a.map + a.map0[s];
      const uint32_t A = a.n_alleles[s];
      for (uint32_t al = 0; al < A; ++al) {
        double v = -__builtin_huge_val();
        for (uint32_t h = 0; h < H; ++h) {
          if (m[h] != (int16_t)al) continue;
          double w = row[h];
          w = w < floor_v ? floor_v : w;
          if (w > v) v = w;
        }
        a.out[a.allele0[s] + (uint64_t)al * g.n_reads + x] = v;
      }
    }
  }
}

// Host-to-device copy of a small staged image by a kernel that reads pinned
// host memory (16 bytes per lane, then the tail bytes).
__global__ void k_copy(char* dst, const char* src, uint64_t n) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t done = 0;
  if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0) {
    const uint64_t n16 = n >> 4;
    for (uint64_t i = tid; i < n16; i += stride)
      reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
    done = n16 << 4;
  }
  for (uint64_t i = done + tid; i < n; i += stride) dst[i] = src[i];
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Results start as NaN with no flags (pairs with an empty side keep that):
// n doubles, then n flag bytes written as ceil(n / 8) zero words.
__global__ void k_init_results(uint64_t* out, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t* flags = out + n;
  const uint64_t nf = (n + 7) / 8;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    out[i] = 0x7ff8000000000000ull;
    if (i < nf) flags[i] = 0;
  }
}

using Fp32Fn = void (*)(Fp32Args);
using Fp64Fn = void (*)(Fp64Args);

template <int W, bool XF, int OCC, int... R>
constexpr auto fp32_row(std::integer_sequence<int, R...>) {
  return std::array<Fp32Fn, sizeof...(R)>{{&k_fp32<W, R + 1, XF, OCC>...}};
}
template <int W, int OCC, int... R>
constexpr auto fp32pt_row(std::integer_sequence<int, R...>) {
  return std::array<Fp32Fn, sizeof...(R)>{{&k_fp32pt<W, R + 1, true, OCC>...}};
}

// FP32 variants: W = 16 with 1..16 rows per lane at occupancy targets 2, 3
// and 4 waves per SIMD; W = 32 (1..16) and W = 64 (1..12) at 2 waves; each
// in the plain and in the scaled form.
using Row16 = std::array<Fp32Fn, 16>;
const Row16 kFp32W16[2][3] = {
    {fp32_row<16, false, 2>(std::make_integer_sequence<int, 16>{}),
     fp32_row<16, false, 3>(std::make_integer_sequence<int, 16>{}),
     fp32_row<16, false, 4>(std::make_integer_sequence<int, 16>{})},
    {fp32_row<16, true, 2>(std::make_integer_sequence<int, 16>{}),
     fp32_row<16, true, 3>(std::make_integer_sequence<int, 16>{}),
     fp32_row<16, true, 4>(std::make_integer_sequence<int, 16>{})}};
const Row16 kFp32W32[2] = {fp32_row<32, false, 2>(std::make_integer_sequence<int, 16>{}),
                           fp32_row<32, true, 2>(std::make_integer_sequence<int, 16>{})};
const std::array<Fp32Fn, 13> kFp32W64[2] = {fp32_row<64, false, 2>(std::make_integer_sequence<int, 13>{}),
                                            fp32_row<64, true, 2>(std::make_integer_sequence<int, 13>{})};
const Row16 kFp32W8[2] = {fp32_row<8, false, 2>(std::make_integer_sequence<int, 16>{}),
                          fp32_row<8, true, 2>(std::make_integer_sequence<int, 16>{})};
// W = 8 with 17-19 rows per lane (reads of 135-151 bases in 152 rows): the
// compare kernel spills at these sizes and is there only for batches that
// cannot take the prior-table kernel (gap qualities outside the scaled form)
const std::array<Fp32Fn, 3> kFp32W8Long[2] = {{{&k_fp32<8, 17, false, 2>, &k_fp32<8, 18, false, 2>,
                                                &k_fp32<8, 19, false, 2>}},
                                              {{&k_fp32<8, 17, true, 2>, &k_fp32<8, 18, true, 2>,
                                                &k_fp32<8, 19, true, 2>}}};

const std::array<Fp64Fn, kNumFp64Bins> kFp64[2] = {
    {{&k_fp64<16, 4, false>, &k_fp64<16, 8, false>, &k_fp64<32, 5, false>, &k_fp64<32, 8, false>,
      &k_fp64<64, 12, false>}},
    {{&k_fp64<16, 4, true>, &k_fp64<16, 8, true>, &k_fp64<32, 5, true>, &k_fp64<32, 8, true>,
      &k_fp64<64, 12, true>}}};
const std::array<Fp64Fn, kNumFp64Bins> kFp64Pt = {
    {&k_fp64pt<16, 4>, &k_fp64pt<16, 8>, &k_fp64pt<32, 5>, &k_fp64pt<32, 8>, &k_fp64pt<64, 12>}};

// Prior-table variants (scaled form only): W = 8, 16 and 32 at 2, 3, 4 waves
// per SIMD, W = 64 at 2.
const Row16 kFp32PtW16[3] = {fp32pt_row<16, 2>(std::make_integer_sequence<int, 16>{}),
                             fp32pt_row<16, 3>(std::make_integer_sequence<int, 16>{}),
                             fp32pt_row<16, 4>(std::make_integer_sequence<int, 16>{})};
const Row16 kFp32PtW32[3] = {fp32pt_row<32, 2>(std::make_integer_sequence<int, 16>{}),
                             fp32pt_row<32, 3>(std::make_integer_sequence<int, 16>{}),
                             fp32pt_row<32, 4>(std::make_integer_sequence<int, 16>{})};
const std::array<Fp32Fn, 13> kFp32PtW64 = fp32pt_row<64, 2>(std::make_integer_sequence<int, 13>{});
const Row16 kFp32PtW8[3] = {fp32pt_row<8, 2>(std::make_integer_sequence<int, 16>{}),
                            fp32pt_row<8, 3>(std::make_integer_sequence<int, 16>{}),
                            fp32pt_row<8, 4>(std::make_integer_sequence<int, 16>{})};
// 8 lanes x 17-19 rows at 2 waves per SIMD (19 rows: 249 VGPRs, no spill;
// the 20 KB prior table leaves room for 7 waves per CU)
const std::array<Fp32Fn, 3> kFp32PtW8Long = {{&k_fp32pt<8, 17, true, 2>, &k_fp32pt<8, 18, true, 2>,
                                              &k_fp32pt<8, 19, true, 2>}};

Fp32Fn fp32pt_fn(int W, int RPL, int occ) {
  if (RPL < 1) return nullptr;
  const int o = std::min(4, std::max(2, occ)) - 2;
  if (W == 16 && RPL <= 16) return kFp32PtW16[o][RPL - 1];
  if (W == 32 && RPL <= 16) return kFp32PtW32[o][RPL - 1];
  if (W == 64 && RPL <= 13) return kFp32PtW64[RPL - 1];
  if (W == 8 && RPL <= 16) return kFp32PtW8[o][RPL - 1];
  if (W == 8 && RPL <= 19) return kFp32PtW8Long[RPL - 17];
  return nullptr;
}

Fp32Fn fp32_fn(int W, int RPL, bool xf, int occ, bool pt) {
  if (pt) return xf ? fp32pt_fn(W, RPL, occ) : nullptr;
  if (RPL < 1) return nullptr;
  if (W == 16 && RPL <= 16) return kFp32W16[xf][std::min(4, std::max(2, occ)) - 2][RPL - 1];
  if (W == 32 && RPL <= 16) return kFp32W32[xf][RPL - 1];
  if (W == 64 && RPL <= 13) return kFp32W64[xf][RPL - 1];
  if (W == 8 && RPL <= 16) return kFp32W8[xf][RPL - 1];
  if (W == 8 && RPL <= 19) return kFp32W8Long[xf][RPL - 17];
  return nullptr;
}
}  // namespace

bool fp32_supported(int W, int RPL) { return fp32_fn(W, RPL, false, 2, false) != nullptr; }

hipError_t launch_copy(void* dst, const void* src, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t n16 = (n + 15) >> 4;
  const unsigned blocks = (unsigned)std::min<uint64_t>(1024, std::max<uint64_t>(1, (n16 + 255) / 256));
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, stream, static_cast<char*>(dst),
                     static_cast<const char*>(src), n);
  return hipGetLastError();
}

hipError_t launch_fp32(int W, int RPL, bool xf, int occ, bool pt, int grid, hipStream_t stream, const Fp32Args& a) {
  Fp32Fn f = fp32_fn(W, RPL, xf, occ, pt);
  if (!f) return hipErrorInvalidValue;
  hipLaunchKernelGGL(f, dim3(grid), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_fp64(int bin, bool pt, bool scaled, int grid, hipStream_t stream, const Fp64Args& a) {
  if (bin < 0 || bin >= kNumFp64Bins || (pt && !scaled)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pt ? kFp64Pt[bin] : kFp64[scaled][bin], dim3(grid), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_prep(const PrepArgs& a, hipStream_t stream) {
  if (a.n_reads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep, dim3(a.n_reads), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

hipError_t launch_marginalize(const MargArgs& a, hipStream_t stream) {
  if (a.n_regions == 0) return hipSuccess;
  hipLaunchKernelGGL(k_marginalize, dim3(a.n_regions), dim3(256), 0, stream, a);
  return hipGetLastError();
}

// Counters to zero and results to NaN with flags 0, one grid-stride pass.
__global__ void k_begin(uint32_t* p, uint32_t n, uint64_t* out, uint64_t n_out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = i0; i < n; i += stride) p[i] = 0;
  uint64_t* flags = out + n_out;
  const uint64_t nf = (n_out + 7) / 8;
  for (uint64_t i = i0; i < n_out; i += stride) {
    out[i] = 0x7ff8000000000000ull;
    if (i < nf) flags[i] = 0;
  }
}

hipError_t launch_begin(uint32_t* p, uint32_t n, double* out, uint64_t n_out, hipStream_t stream) {
  const uint64_t items = std::max<uint64_t>(n, n_out);
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((items + 255) / 256, 4096));
  hipLaunchKernelGGL(k_begin, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, reinterpret_cast<uint64_t*>(out),
                     n_out);
  return hipGetLastError();
}

hipError_t launch_init_results(double* out, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_init_results, dim3((unsigned)blocks), dim3(256), 0, stream, reinterpret_cast<uint64_t*>(out), n);
  return hipGetLastError();
}

}  // namespace rvc

namespace rvc {

// Resident 64-lane blocks per CU for a kernel (register and LDS limited).
// The occupancy queries are cached per kernel. Serving threads and callers of
// the one-region path ask concurrently: every slot is an atomic, and two
// threads that both find it empty store the same value.
int fp32_blocks_per_cu(int W, int RPL, bool xf, int occ, bool pt) {
  static std::atomic<int> cache[2][2][3][4][kMaxRpl + 1] = {};
  const int w = W == 16 ? 0 : W == 32 ? 1 : W == 64 ? 2 : 3;
  const int o = std::min(4, std::max(2, occ)) - 2;
  if (RPL < 1 || RPL > kMaxRpl) return 1;
  std::atomic<int>& slot = cache[pt][xf][o][w][RPL];
  int n = slot.load(std::memory_order_relaxed);
  if (n == 0) {
    Fp32Fn f = fp32_fn(W, RPL, xf, occ, pt);
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64, 0) != hipSuccess || n < 1) n = 1;
    slot.store(n, std::memory_order_relaxed);
  }
  return n;
}

int fp64_blocks_per_cu(int bin, bool pt, bool scaled) {
  static std::atomic<int> cache[2][2][kNumFp64Bins] = {};
  if (bin < 0 || bin >= kNumFp64Bins) return 1;
  std::atomic<int>& slot = cache[pt][scaled][bin];
  int n = slot.load(std::memory_order_relaxed);
  if (n == 0) {
    Fp64Fn f = pt ? kFp64Pt[bin] : kFp64[scaled][bin];
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64, 0) != hipSuccess || n < 1) n = 1;
    slot.store(n, std::memory_order_relaxed);
  }
  return n;
}

}  // namespace rvc
